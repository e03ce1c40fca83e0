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
 * steps): on host threads, one survivor at a time per thread, or on the GPU
 * (end of this file), one wave per survivor.
 */
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "e13.h"
#include "e13_ctx.h"

#define DEEP_LEFT 4096         /* cells left of cell 0 (deep.c:13) */
#define DEEP_BIG_LEFT (1 << 22) /* deep_big.c:14: cells each side beyond the block */

/* The tape of one run: deep.c's, or deep_big.c's 4M cells each side. */
static long long deep_left(int big) { return big ? DEEP_BIG_LEFT : DEEP_LEFT; }
static long long deep_size(int nmax, int big) {
  return big ? DEEP_BIG_LEFT + (1LL << nmax) + DEEP_BIG_LEFT + 4096 : DEEP_LEFT + (1LL << nmax) + 4096;
}

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

/* One survivor (number id) to stage nmax on tape[0 .. deep_size).  big:
   deep_big.c's tape and its report of head, state and ones at the limit. */
static int deep_run(const deep_machine *m, int id, int nmax, int big, uint8_t *tape, char *out, int len) {
  const long long size = deep_size(nmax, big), left = deep_left(big);
  memset(tape, 0, (size_t)size);
  uint8_t *t0 = tape + left; /* cell k at t0[k] */
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
        if (big) { /* deep_big.c:34: where the machine is and the extent of its ones */
          long long lo1 = 0, hi1 = 0, cnt = 0;
          for (long long k = -left + 1; k < size - left; k++)
            if (t0[k]) {
              if (!cnt) lo1 = k;
              hi1 = k;
              cnt++;
            }
          if (appendf(out, len, &pos, " head=%lld state=%c ones=%lld in [%lld,%lld]", head, (char)(65 + st), cnt, lo1,
                      hi1))
            return -1;
        }
        failed = 1;
        break;
      }
      if (head < -left + 2 || head > blk + (big ? DEEP_BIG_LEFT : 2000)) {
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
    for (long long k = -3; k >= -left + 1; k--)
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

int e13_deep_check_line(const char *line, int id, int nmax, int big, char *out, int outlen) {
  if (!line || !out || outlen < 2 || nmax < 1 || nmax > 24) return E13_EINVAL;
  deep_machine m;
  if (!deep_parse(line, &m)) return E13_EINVAL;
  uint8_t *tape = (uint8_t *)malloc((size_t)deep_size(nmax, big));
  if (!tape) return E13_ENOMEM;
  const int r = deep_run(&m, id, nmax, big, tape, out, outlen);
  free(tape);
  return r < 0 ? E13_EINVAL : r;
}

/* Many survivors on host threads: every line of `text` with table= and st0=
   is one survivor, numbered from 1 in input order; the result lines go to out
   in the same order. */
typedef struct {
  const char **lines;
  int n, nmax, recheck, big; /* recheck: the counter re-check instead of the deep check; big: deep_big.c */
  char *res;            /* n x DEEP_OUT */
  int next, err;
  pthread_mutex_t mu;
} deep_job;

#define DEEP_OUT 20480 /* a result line (a counter re-check line holds two sets of up to 8000 bytes) */

static int rc_run(const deep_machine *m, int id, int nmax, uint8_t *tape, char (*ctr)[8192], char *out, int len);
static int deep_parse_recheck(const char *line, deep_machine *m);

static void *deep_worker(void *arg) {
  deep_job *j = (deep_job *)arg;
  uint8_t *tape = (uint8_t *)malloc(j->recheck ? (size_t)(2 * 4000000LL + 1) : (size_t)deep_size(j->nmax, j->big));
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
                                             deep_run(&m, i + 1, j->nmax, j->big, tape, o, DEEP_OUT) >= 0);
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

static int deep_many(const char *text, int nmax, int threads, char *out, int64_t outlen, int recheck, int big) {
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
  j.big = big;
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

int e13_deep_check(const char *text, int nmax, int big, int threads, char *out, int64_t outlen) {
  if (!text || !out || outlen < 1 || nmax < 1 || nmax > 24) return E13_EINVAL;
  return deep_many(text, nmax, threads, out, outlen, 0, big);
}

int e13_recheck_counter(const char *text, int nmax, int threads, char *out, int64_t outlen) {
  if (!text || !out || outlen < 1 || nmax < 1 || nmax > 19) return E13_EINVAL;
  return deep_many(text, nmax, threads, out, outlen, 1, 0);
}

/* ------------------------------------------------------------------------ */
/* Classification of deep-check lines (the reference's src/classify_deep.py):*/
/* per survivor, the first counter change within a class c(n), counters     */
/* equal across classes, the failure the line ends with (times and places   */
/* removed), or a candidate; the kinds counted apart for absolute and       */
/* relative starts, most frequent first, then an example survivor line for  */
/* every candidate or free-cell kind.                                        */
/* ------------------------------------------------------------------------ */
typedef struct {
  char kind[512];
  int rel, count;
  const char *example; /* the first survivor line of this kind */
} cl_kind;

static int cl_c(int n) { return n <= 2 ? 0 : (n <= 8 ? 1 : 2); }

/* The failure text with "(t=N)", "(limit, t=N)" and "at N" removed, as the
   reference's regular expression does, scanning left to right. */
static void cl_strip(const char *a, size_t n, char *o, size_t cap) {
  size_t j = 0;
  for (size_t i = 0; i < n && j + 1 < cap;) {
    size_t k = 0;
    if (!strncmp(a + i, "(t=", 3) || !strncmp(a + i, "(limit, t=", 10)) {
      size_t p = i + (a[i + 1] == 't' ? 3 : 10), q = p;
      while (q < n && a[q] >= '0' && a[q] <= '9') q++;
      if (q > p && q < n && a[q] == ')') k = q + 1 - i;
    } else if (!strncmp(a + i, "at ", 3)) {
      size_t p = i + 3;
      if (p < n && a[p] == '-') p++;
      size_t q = p;
      while (q < n && a[q] >= '0' && a[q] <= '9') q++;
      if (q > p) k = q - i;
    }
    if (k) i += k;
    else o[j++] = a[i++];
  }
  o[j] = 0;
}

static char *cl_line(const char **p) { /* the next line of a text, or NULL */
  if (!**p) return NULL;
  const char *e = strchr(*p, '\n');
  const size_t n = e ? (size_t)(e - *p) : strlen(*p);
  char *l = (char *)malloc(n + 1);
  if (l) {
    memcpy(l, *p, n);
    l[n] = 0;
  }
  *p += n + (e ? 1 : 0);
  return l;
}

int e13_classify_deep(const char *survivors, const char *deep, char *out, int64_t outlen) {
  if (!survivors || !deep || !out || outlen < 1) return E13_EINVAL;
  int cap = 4096, ns = 0, nd = 0, nk = 0, rc = E13_OK;
  char **sv = (char **)malloc(sizeof(char *) * (size_t)cap), **dv = (char **)malloc(sizeof(char *) * (size_t)cap);
  cl_kind *kinds = (cl_kind *)calloc(256, sizeof(cl_kind));
  if (!sv || !dv || !kinds) rc = E13_ENOMEM;
  for (const char *p = survivors; !rc;) { /* lines with table= */
    char *l = cl_line(&p);
    if (!l) break;
    if (!strstr(l, "table=")) {
      free(l);
      continue;
    }
    if (ns == cap) rc = E13_EINVAL;
    else sv[ns++] = l;
  }
  for (const char *p = deep; !rc;) { /* lines that start with "survivor" */
    char *l = cl_line(&p);
    if (!l) break;
    if (strncmp(l, "survivor", 8)) {
      free(l);
      continue;
    }
    if (nd == cap) rc = E13_EINVAL;
    else dv[nd++] = l;
  }
  if (!rc && ns != nd) rc = E13_EINVAL;
  for (int i = 0; !rc && i < nd; i++) {
    const char *d = dv[i], *body = strstr(d, "):");
    if (!body) {
      rc = E13_EINVAL;
      break;
    }
    body += 2;
    const char *bar = strchr(body, '|');
    const size_t slen = bar ? (size_t)(bar - body) : strlen(body);
    /* n -> its "cnt@lo" text, from " nK:VALUE" items before the '|' */
    char val[64][64];
    int have[64] = {0}, nmax = 0;
    for (const char *q = body; q < body + slen; q++) {
      if (*q != 'n' || q[1] < '0' || q[1] > '9') continue;
      char *e;
      const long n = strtol(q + 1, &e, 10);
      if (*e != ':' || n < 1 || n > 63) continue;
      const char *v = e + 1;
      size_t k = 0;
      while (v + k < body + slen && v[k] != ' ' && v[k] != '\t' && k < 63) k++;
      if (!k) continue;
      memcpy(val[n], v, k);
      val[n][k] = 0;
      have[n] = 1;
      if (n > nmax) nmax = (int)n;
      q = v + k - 1;
    }
    char kind[512];
    kind[0] = 0;
    int first[3] = {0, 0, 0};
    for (int n = 1; n <= nmax && !kind[0]; n++) {
      if (!have[n]) continue;
      const int c = cl_c(n);
      if (first[c] && strcmp(val[first[c]], val[n]))
        snprintf(kind, sizeof kind, "counter changes at n=%d (%s -> %s)", n, val[first[c]], val[n]);
      else if (!first[c])
        first[c] = n;
    }
    if (!kind[0]) {
      for (int a = 0; a < 3; a++)
        for (int b = a + 1; b < 3; b++)
          if (first[a] && first[b] && !strcmp(val[first[a]], val[first[b]]))
            snprintf(kind, sizeof kind, "counter equal for different c");
    }
    if (!kind[0] && bar) { /* the failure the line ends with */
      const char *a = bar + 1, *b = d + strlen(d);
      while (a < b && (*a == ' ' || *a == '\t')) a++;
      while (b > a && (b[-1] == ' ' || b[-1] == '\t' || b[-1] == '\r')) b--;
      if (b > a) cl_strip(a, (size_t)(b - a), kind, sizeof kind);
    }
    if (!kind[0]) snprintf(kind, sizeof kind, "CANDIDATE (consistent to n=%d)", nmax);
    const int rel = strstr(d, " rel)") != NULL;
    int j = 0;
    while (j < nk && (kinds[j].rel != rel || strcmp(kinds[j].kind, kind))) j++;
    if (j == nk) {
      if (nk == 256) {
        rc = E13_EINVAL;
        break;
      }
      snprintf(kinds[nk].kind, sizeof kinds[nk].kind, "%s", kind);
      kinds[nk].rel = rel;
      kinds[nk].example = sv[i];
      nk++;
    }
    kinds[j].count++;
  }
  int64_t pos = 0;
  int done[256] = {0};
  for (int r = 0; !rc && r < nk; r++) { /* most frequent first; ties in order of first appearance */
    int best = -1;
    for (int j = 0; j < nk; j++)
      if (!done[j] && (best < 0 || kinds[j].count > kinds[best].count)) best = j;
    done[best] = 1;
    const int k = snprintf(out + pos, (size_t)(outlen - pos), "%6d  %s  %s\n", kinds[best].count,
                           kinds[best].rel ? "rel" : "abs", kinds[best].kind);
    if (k < 0 || pos + k >= outlen) rc = E13_EINVAL;
    else pos += k;
  }
  for (int j = 0; !rc && j < nk; j++) {
    if (strncmp(kinds[j].kind, "CANDIDATE", 9) && !strstr(kinds[j].kind, "FREE")) continue;
    const char q = strchr(kinds[j].kind, '\'') ? '"' : '\''; /* Python's repr of the (kind, rel) pair */
    const int k = snprintf(out + pos, (size_t)(outlen - pos), "example (%c%s%c, '%s') %s\n", q, kinds[j].kind, q,
                           kinds[j].rel ? "rel" : "abs", kinds[j].example);
    if (k < 0 || pos + k >= outlen) rc = E13_EINVAL;
    else pos += k;
  }
  if (!rc) out[pos] = 0;
  for (int i = 0; i < ns; i++) free(sv[i]);
  for (int i = 0; i < nd; i++) free(dv[i]);
  free(sv);
  free(dv);
  free(kinds);
  return rc ? rc : nd;
}

/* ------------------------------------------------------------------------ */
/* The deep check (deep.c, 4096 cells left of cell 0) and the counter        */
/* re-check (recheck.c) on the GPU: one wave per survivor, its tape in LDS   */
/* (k_deep, e13_deepk.hip), the same lines as the host versions above.  A    */
/* re-check whose machine leaves the tape held on the GPU (recheck.c's tape  */
/* is 4e6 cells each side) runs on the host instead; so does a table whose   */
/* moves name a state it does not have.                                      */
/* ------------------------------------------------------------------------ */
int e13k_launch_deep(const E13DeepArgs *a, int nblocks, hipStream_t s);

#define DG_BUDGET (1LL << 20) /* loop iterations per survivor per launch */

/* deep.c's counter at a stage start from the snapshot of the tape's left side
   (bits 0 .. left-1): the ones at cells -3 down to -left+1, and the leftmost. */
static void dg_count(const uint32_t *sn, int left, long long *cnt, long long *lo) {
  *cnt = 0;
  *lo = 0;
  for (int q = left - 3; q >= 1; q--)
    if ((sn[q >> 5] >> (q & 31)) & 1u) {
      ++*cnt;
      *lo = q - left;
    }
}

/* recheck.c's counter set text (rc_counter) from the snapshot: cells left of
   the leftmost head position are blank, so listing every one down to the
   tape's end is the same list. */
static void dg_set(const uint32_t *sn, int left, char *q) {
  int len = 0;
  q[0] = 0;
  for (int b = left - 3; b >= 0 && len < 8000; b--)
    if ((sn[b >> 5] >> (b & 31)) & 1u) len += snprintf(q + len, (size_t)(RC_SET - len), "%d,", b - left);
}

static int deep_gpu(e13_ctx *c, const char *text, int nmax, int recheck, char *out, int64_t outlen) {
  if (!c || !text || !out || outlen < 1 || nmax < 1 || nmax > 18) return E13_EINVAL;
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
  const int left = DEEP_LEFT;
  const int W = (int)((left + (1LL << nmax) + 4096 + 31) / 32), SW = left / 32;
  deep_machine *ms = lines && n ? (deep_machine *)calloc((size_t)n, sizeof *ms) : NULL;
  uint32_t *rows = n ? (uint32_t *)calloc((size_t)n * E13_DEEP_MAXQ, 4) : NULL;
  int32_t *param = n ? (int32_t *)calloc((size_t)n * 4, 4) : NULL;
  E13DeepState *st = n ? (E13DeepState *)calloc((size_t)n, sizeof *st) : NULL;
  int32_t *list = n ? (int32_t *)malloc((size_t)n * 4) : NULL;
  uint32_t *snap = n ? (uint32_t *)malloc((size_t)n * (size_t)(nmax + 1) * SW * 4) : NULL;
  char *res = n ? (char *)calloc((size_t)n, DEEP_OUT) : NULL;
  int rc = lines && (!n || (ms && rows && param && st && list && snap && res)) ? E13_OK : E13_ENOMEM;
  const long long kmul = recheck ? 8 : 64;
  const int timing = getenv("E13_TIMING") && *getenv("E13_TIMING") != '0'; /* a line per launch on stderr */
  for (int i = 0; !rc && i < n; i++) {
    deep_machine *m = &ms[i];
    if (!(recheck ? deep_parse_recheck(lines[i], m) : deep_parse(lines[i], m))) rc = E13_EINVAL;
    int host = 0; /* a move to a state the table does not have: the host's loop */
    for (int q = 0; !rc && q < E13_DEEP_MAXQ; q++) {
      uint32_t ent[2];
      for (int a = 0; a < 2; a++) {
        if (q >= m->nq || m->stop[q][a]) {
          ent[a] = 0x8000u | ((q >= m->nq ? 1u : (uint32_t)m->stop[q][a]) << 13);
        } else {
          if (m->nx[q][a] < 0 || m->nx[q][a] >= m->nq) host = 1;
          ent[a] = (uint32_t)m->wr[q][a] | ((m->dir[q][a] > 0 ? 1u : 0u) << 1) | ((uint32_t)(m->nx[q][a] & 31) << 8);
        }
      }
      rows[(size_t)i * E13_DEEP_MAXQ + q] = ent[0] | (ent[1] << 16);
    }
    if (m->st0 < 0 || m->st0 >= E13_DEEP_MAXQ || m->head0 < -left + 2 || m->head0 > 4096) host = 1;
    param[4 * i + 0] = m->st0;
    param[4 * i + 1] = (int32_t)(m->head0 + left);
    param[4 * i + 2] = m->rel;
    st[i].q = left;
    st[i].minq = left;
    st[i].n = 1;
    st[i].limit = kmul * 4 + 100000;
    st[i].status = host ? 2 : 0;
  }
  /* device buffers, launches until every survivor is finished */
  uint32_t *d_rows = NULL, *d_tape = NULL, *d_snap = NULL;
  int32_t *d_param = NULL, *d_list = NULL;
  E13DeepState *d_st = NULL;
  float kms = 0;
  int launches = 0;
  if (!rc && n) {
    if (hipSetDevice(c->device) != hipSuccess) rc = E13_EHIP;
    if (!rc && (hipMalloc((void **)&d_rows, (size_t)n * E13_DEEP_MAXQ * 4) != hipSuccess ||
                hipMalloc((void **)&d_param, (size_t)n * 16) != hipSuccess ||
                hipMalloc((void **)&d_st, (size_t)n * sizeof(E13DeepState)) != hipSuccess ||
                hipMalloc((void **)&d_list, (size_t)n * 4) != hipSuccess ||
                hipMalloc((void **)&d_tape, (size_t)n * (size_t)W * 4) != hipSuccess ||
                hipMalloc((void **)&d_snap, (size_t)n * (size_t)(nmax + 1) * SW * 4) != hipSuccess))
      rc = E13_ENOMEM;
    if (!rc && (hipMemcpy(d_rows, rows, (size_t)n * E13_DEEP_MAXQ * 4, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(d_param, param, (size_t)n * 16, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(d_st, st, (size_t)n * sizeof(E13DeepState), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemset(d_tape, 0, (size_t)n * (size_t)W * 4) != hipSuccess ||
                hipMemset(d_snap, 0, (size_t)n * (size_t)(nmax + 1) * SW * 4) != hipSuccess))
      rc = E13_EHIP;
    E13DeepArgs a;
    memset(&a, 0, sizeof a);
    a.rows = d_rows;
    a.param = d_param;
    a.state = d_st;
    a.tape = d_tape;
    a.snap = d_snap;
    a.list = d_list;
    a.nmax = nmax;
    a.W = W;
    a.left = left;
    a.mode = recheck;
    a.kmul = kmul;
    a.budget = DG_BUDGET;
    for (;;) {
      if (rc) break;
      int m = 0;
      for (int i = 0; i < n; i++)
        if (st[i].status == 0) list[m++] = i;
      if (m == 0) break;
      if (hipMemcpyAsync(d_list, list, (size_t)m * 4, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
          hipEventRecord(c->ev0, c->stream) != hipSuccess || e13k_launch_deep(&a, m, c->stream) != 0 ||
          hipEventRecord(c->ev1, c->stream) != hipSuccess ||
          hipMemcpyAsync(st, d_st, (size_t)n * sizeof(E13DeepState), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
          hipStreamSynchronize(c->stream) != hipSuccess) {
        rc = E13_EHIP;
        break;
      }
      float ms1 = 0;
      if (hipEventElapsedTime(&ms1, c->ev0, c->ev1) == hipSuccess) kms += ms1;
      launches++;
      if (timing) fprintf(stderr, "libe13: deep launch %d: %d survivors, %.1f ms\n", launches, m, ms1);
      for (int i = 0; i < n; i++)
        if (st[i].status == 3) rc = E13_ELDS;
    }
    if (!rc && hipMemcpy(snap, d_snap, (size_t)n * (size_t)(nmax + 1) * SW * 4, hipMemcpyDeviceToHost) != hipSuccess)
      rc = E13_EHIP;
    c->stats.kernel_ms += kms;
    c->stats.launches += launches;
  }
  hipFree(d_rows);
  hipFree(d_param);
  hipFree(d_st);
  hipFree(d_list);
  hipFree(d_tape);
  hipFree(d_snap);
  /* the result lines, as deep_run / rc_run write them */
  uint8_t *htape = NULL;
  char (*ctr)[RC_SET] = recheck && n ? (char (*)[RC_SET])calloc(20, RC_SET) : NULL;
  if (!rc && recheck && n && !ctr) rc = E13_ENOMEM;
  for (int i = 0; !rc && i < n; i++) {
    const deep_machine *m = &ms[i];
    char *o = res + (size_t)i * DEEP_OUT;
    const E13DeepState *s = &st[i];
    const uint32_t *sn = snap + (size_t)i * (size_t)(nmax + 1) * SW;
    const int reached = s->n - 1; /* stages whose start was found */
    if (s->status == 2) { /* the host's loop */
      if (!htape) htape = (uint8_t *)malloc(recheck ? (size_t)(2 * RC_SIDE + 1) : (size_t)deep_size(nmax, 0));
      if (!htape) rc = E13_ENOMEM;
      else if ((recheck ? rc_run(m, i + 1, nmax, htape, ctr, o, DEEP_OUT) : deep_run(m, i + 1, nmax, 0, htape, o, DEEP_OUT)) < 0)
        rc = E13_EINVAL;
      continue;
    }
    int pos = 0;
    if (!recheck) {
      if (appendf(o, DEEP_OUT, &pos, "survivor %d (nq=%d st0=%d%s):", i + 1, m->nq, m->st0, m->rel ? " rel" : "")) rc = E13_EINVAL;
      for (int k = 1; !rc && k <= reached; k++) {
        long long cnt, lo;
        dg_count(sn + (size_t)k * SW, left, &cnt, &lo);
        if (appendf(o, DEEP_OUT, &pos, " n%d:%lld", k, cnt) || (cnt && appendf(o, DEEP_OUT, &pos, "@%lld", lo)))
          rc = E13_EINVAL;
      }
      if (!rc && s->fail) {
        const int r = s->fail == 1   ? appendf(o, DEEP_OUT, &pos, " | stage %d NOT reached (limit, t=%lld)", s->fail_n,
                                               (long long)s->fail_t)
                      : s->fail == 2 ? appendf(o, DEEP_OUT, &pos, " | stage %d: off tape at %lld", s->fail_n,
                                               (long long)s->fail_q - left)
                      : s->fail == 3 ? appendf(o, DEEP_OUT, &pos, " | FREE cell %c%d read before stage %d (t=%lld)",
                                               'A' + s->fail_st, s->fail_sym, s->fail_n, (long long)s->fail_t)
                                     : appendf(o, DEEP_OUT, &pos, " | HALT before stage %d (t=%lld)", s->fail_n,
                                               (long long)s->fail_t);
        if (r) rc = E13_EINVAL;
      }
      if (!rc && appendf(o, DEEP_OUT, &pos, "\n")) rc = E13_EINVAL;
      continue;
    }
    /* the counter re-check's verdict from the counter sets at the starts found */
    char verdict[2 * RC_SET + 128];
    verdict[0] = 0;
    int failed = 0;
    for (int k = 1; k <= reached && !failed; k++) {
      dg_set(sn + (size_t)k * SW, left, ctr[k]);
      const int cc = k <= 2 ? 0 : (k <= 8 ? 1 : 2), first = cc == 0 ? 1 : (cc == 1 ? 3 : 9);
      if (k != first && strcmp(ctr[k], ctr[first])) {
        snprintf(verdict, sizeof verdict, "counter changes at n=%d (%s -> %s)", k, ctr[first][0] ? ctr[first] : "none",
                 ctr[k][0] ? ctr[k] : "none");
        failed = 1;
      }
    }
    if (!failed && s->fail) {
      failed = 1;
      if (s->fail == 1) snprintf(verdict, sizeof verdict, "stage %d not reached (limit)", s->fail_n);
      else if (s->fail == 3)
        snprintf(verdict, sizeof verdict, "free cell %c%d read before stage %d", 'A' + s->fail_st, s->fail_sym, s->fail_n);
      else if (s->fail == 4) snprintf(verdict, sizeof verdict, "halt before stage %d", s->fail_n);
      else rc = E13_EINVAL; /* off the tape in the re-check is the host's case */
    }
    if (!rc && !failed) {
      if (nmax >= 9 && (!strcmp(ctr[1], ctr[3]) || !strcmp(ctr[3], ctr[9]) || !strcmp(ctr[1], ctr[9])))
        snprintf(verdict, sizeof verdict, "counter equal for different c");
      else
        snprintf(verdict, sizeof verdict, "CONSISTENT to n=%d (c2 counter %s)", nmax, nmax >= 9 ? ctr[9] : "");
    }
    if (!rc && appendf(o, DEEP_OUT, &pos, "%d: %s\n", i + 1, verdict)) rc = E13_EINVAL;
  }
  free(htape);
  free(ctr);
  int64_t pos = 0;
  for (int i = 0; !rc && i < n; i++) {
    const size_t k = strlen(res + (size_t)i * DEEP_OUT);
    if (pos + (int64_t)k >= outlen) {
      rc = E13_EINVAL;
      break;
    }
    memcpy(out + pos, res + (size_t)i * DEEP_OUT, k);
    pos += (int64_t)k;
  }
  out[pos < outlen ? pos : outlen - 1] = 0;
  free(ms);
  free(rows);
  free(param);
  free(st);
  free(list);
  free(snap);
  free(res);
  free(lines);
  free(copy);
  return rc ? rc : n;
}

int e13_deep_check_gpu(e13_ctx *c, const char *text, int nmax, char *out, int64_t outlen) {
  return deep_gpu(c, text, nmax, 0, out, outlen);
}

int e13_recheck_counter_gpu(e13_ctx *c, const char *text, int nmax, char *out, int64_t outlen) {
  return deep_gpu(c, text, nmax, 1, out, outlen);
}
