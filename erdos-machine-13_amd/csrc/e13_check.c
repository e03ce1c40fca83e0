/* e13_check.c -- the independent survivor check with crafted stage starts
 * (the reference's src/check.py), restated in C with the same printed lines.
 *
 * For a survivor table: (1) run from the blank tape and find stage starts
 * n = 1..11 (the ones at cells >= -2 exactly the block -1 .. 2^n - 2; from
 * stage 2 on also the first start's state, with the head at the same cell or
 * at the same offset from the block end), recording the counter (the ones at
 * cells < -2) at each; the counter must depend only on c(n) = #{1 <= k < n :
 * 2^k has no digit 2 in base 3} and differ between values of c.  (2) From
 * crafted stage starts (the counter of c, a block of m ones, m = 2..40) the
 * machine must halt iff c = 2 and m has no digit 2, and otherwise reach the
 * start with block 2m and the counter of c or c + 1.  With no halt cell in the
 * table, the halt is put into each cell the search never read, in turn.
 * The tape grows as the machine needs (the reference's tape is a set).
 */
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "e13.h"

#define CK_NMAX 11 /* check.py:19 */
#define CK_MMAX 40 /* check.py:20 */

typedef struct {
  int halt; /* the cell halts (None) */
  int w, d, nx;
} ck_cell;

typedef struct { /* a tape of cells k at b[k + off], 0 outside */
  uint8_t *b;
  long long off, cap, ones_ge; /* ones at cells >= -2 */
} ck_tape;

static int ck_fit(ck_tape *t, long long k) {
  if (k + t->off >= 0 && k + t->off < t->cap) return 0;
  long long ncap = t->cap * 2;
  while (ncap < 4 * (llabs(k) + 16)) ncap *= 2;
  uint8_t *nb = (uint8_t *)calloc((size_t)ncap, 1);
  if (!nb) return -1;
  const long long noff = ncap / 2;
  memcpy(nb + (noff - t->off), t->b, (size_t)t->cap);
  free(t->b);
  t->b = nb;
  t->off = noff;
  t->cap = ncap;
  return 0;
}

static int ck_get(const ck_tape *t, long long k) {
  return k + t->off >= 0 && k + t->off < t->cap ? t->b[k + t->off] : 0;
}

static int ck_init(ck_tape *t) {
  t->cap = 1 << 16;
  t->off = t->cap / 2;
  t->ones_ge = 0;
  t->b = (uint8_t *)calloc((size_t)t->cap, 1);
  return t->b ? 0 : -1;
}

static int ck_set(ck_tape *t, long long k, int v) {
  if (ck_fit(t, k)) return -1;
  const int old = t->b[k + t->off];
  t->b[k + t->off] = (uint8_t)v;
  if (k >= -2) t->ones_ge += v - old;
  return 0;
}

static int no_two(long long m) { /* check.py:36-42 */
  for (; m; m /= 3)
    if (m % 3 == 2) return 0;
  return 1;
}

/* The block -1 .. size-2 exactly (check.py:45-48). */
static int ck_block(const ck_tape *t, long long size) {
  if (t->ones_ge != size) return 0;
  for (long long k = -1; k < size - 1; k++)
    if (!ck_get(t, k)) return 0;
  return 1;
}

/* A growable text. */
typedef struct {
  char *s;
  size_t n, cap;
  int err;
} ck_str;

static void sf(ck_str *o, const char *fmt, ...) {
  for (;;) {
    va_list ap;
    va_start(ap, fmt);
    const int k = vsnprintf(o->s + o->n, o->cap - o->n, fmt, ap);
    va_end(ap);
    if (k < 0) {
      o->err = 1;
      return;
    }
    if (o->n + (size_t)k < o->cap) {
      o->n += (size_t)k;
      return;
    }
    char *ns = (char *)realloc(o->s, o->cap * 2 + (size_t)k + 1);
    if (!ns) {
      o->err = 1;
      return;
    }
    o->s = ns;
    o->cap = o->cap * 2 + (size_t)k + 1;
  }
}

/* The counter (ones at cells < -2, ascending) as Python prints a tuple. */
static void ck_counter(const ck_tape *t, ck_str *o) {
  int n = 0;
  sf(o, "(");
  for (long long k = -t->off; k < -2; k++)
    if (ck_get(t, k)) sf(o, n++ ? ", %lld" : "%lld", k);
  sf(o, n == 1 ? ",)" : ")");
}

enum { R_STOP, R_HALT, R_LIMIT };
static const char *res_name[3] = {"stop", "halt", "limit"};

/* run_tracked (check.py:55-77): the stop test before every step.  mode 0:
   stage start `size` (first start: any state and head; later: st0 and the head
   at h0 or h0 - 2 + size); mode 1: crafted target, state st0 and head hh. */
typedef struct {
  int mode, have_first, st0;
  long long size, h0, hh;
} ck_stop;

static int ck_run(const ck_cell rows[][2], ck_tape *t, long long *head, int *st, const ck_stop *sp, long long limit,
                  long long *steps) {
  for (long long i = 0; i < limit; i++) {
    int stop = 0;
    if (ck_block(t, sp->size)) {
      if (sp->mode == 0)
        stop = !sp->have_first || (*st == sp->st0 && (*head == sp->h0 || *head - sp->size == sp->h0 - 2));
      else
        stop = *st == sp->st0 && *head == sp->hh;
    }
    if (stop) {
      *steps = i;
      return R_STOP;
    }
    const int sym = ck_get(t, *head);
    const ck_cell *r = &rows[*st][sym];
    if (r->halt) {
      *steps = i;
      return R_HALT;
    }
    if (r->w != sym && ck_set(t, *head, r->w)) return -1;
    *head += r->d;
    *st = r->nx;
  }
  *steps = limit;
  return R_LIMIT;
}

typedef struct {
  int c;
  char *ctr; /* its counter as printed */
} ck_cnt;

/* One placement of the halt (check.py:103-157): the verdict text after "halt@X: ". */
static int ck_check(const ck_cell rows[][2], int nq, ck_str *o) {
  ck_tape t;
  if (ck_init(&t)) return -1;
  long long head = 0, total = 0, st_h[CK_NMAX + 1], st_t[CK_NMAX + 1];
  int st = 0, st_s[CK_NMAX + 1], rc = 0;
  char *ctrs[CK_NMAX + 1] = {0};
  (void)nq;
  for (int n = 1; n <= CK_NMAX; n++) { /* stages (check.py:80-100) */
    ck_stop sp = {0, n > 1, n > 1 ? st_s[1] : 0, 1LL << n, n > 1 ? st_h[1] : 0, 0};
    long long steps = 0;
    const int r = ck_run(rows, &t, &head, &st, &sp, 40LL * (1LL << (2 * n)) + 1000, &steps);
    if (r < 0) {
      rc = -1;
      goto done;
    }
    total += steps;
    if (r != R_STOP) {
      sf(o, "FAIL stage %d: %s at step %lld", n, res_name[r], total);
      goto done;
    }
    st_s[n] = st;
    st_h[n] = head;
    st_t[n] = total;
    ck_str cs = {(char *)malloc(64), 0, 64, 0};
    if (!cs.s) {
      rc = -1;
      goto done;
    }
    ck_counter(&t, &cs);
    ctrs[n] = cs.s;
  }
  { /* counter consistency (check.py:112-123) */
    ck_cnt cnt[3];
    int ncnt = 0;
    const char *bad = NULL;
    char badbuf[96];
    for (int n = 1; n <= CK_NMAX; n++) {
      int c = 0;
      for (int k = 1; k < n; k++) c += no_two(1LL << k);
      int j = 0;
      while (j < ncnt && cnt[j].c != c) j++;
      if (j < ncnt && strcmp(cnt[j].ctr, ctrs[n])) {
        snprintf(badbuf, sizeof badbuf, "counter differs within c=%d at n=%d", c, n);
        bad = badbuf;
      }
      if (j == ncnt) {
        cnt[ncnt].c = c;
        cnt[ncnt++].ctr = ctrs[n];
      }
    }
    for (int a = 0; a < ncnt; a++)
      for (int b = a + 1; b < ncnt; b++)
        if (!strcmp(cnt[a].ctr, cnt[b].ctr)) bad = "counter equal for different c";
    if (bad) {
      sf(o, "FAIL %s", bad);
      goto done;
    }
    const int s0 = st_s[1];
    const long long h0 = st_h[1];
    int absolute = 1;
    for (int n = 1; n <= CK_NMAX; n++) absolute &= st_h[n] == h0;
    /* crafted stage starts (check.py:127-153) */
    int nfail = 0;
    ck_str fl = {(char *)malloc(256), 0, 256, 0};
    if (!fl.s) {
      rc = -1;
      goto done;
    }
    fl.s[0] = 0;
    for (int c = 0; c <= 2; c++) {
      int jc = 0;
      while (jc < ncnt && cnt[jc].c != c) jc++;
      if (jc == ncnt) continue;
      for (int m = 2; m <= CK_MMAX; m++) {
        ck_tape ct;
        if (ck_init(&ct)) {
          rc = -1;
          break;
        }
        /* the counter of c (its cells from the printed tuple) and the block of m */
        for (const char *p = cnt[jc].ctr; *p;) {
          if (*p == '-' || (*p >= '0' && *p <= '9')) {
            char *e;
            const long long k = strtoll(p, &e, 10);
            if (ck_set(&ct, k, 1)) rc = -1;
            p = e;
          } else {
            p++;
          }
        }
        for (long long k = -1; k < m - 1; k++)
          if (ck_set(&ct, k, 1)) rc = -1;
        long long hd = absolute ? h0 : h0 - 2 + m;
        int cs = s0;
        const long long nxt = 2LL * m;
        const int want_c = no_two(m) ? c + 1 : c, want_halt = no_two(m) && c == 2;
        int jw = 0;
        while (jw < ncnt && cnt[jw].c != want_c) jw++;
        ck_stop sp = {1, 1, s0, nxt, h0, absolute ? h0 : h0 - 2 + nxt};
        long long steps = 0;
        const int r = rc ? -1 : ck_run(rows, &ct, &hd, &cs, &sp, 40LL * m * m + 5000, &steps);
        if (r < 0) {
          rc = -1;
          free(ct.b);
          break;
        }
        ck_str got = {(char *)malloc(64), 0, 64, 0};
        int fail = 0;
        if (got.s) ck_counter(&ct, &got);
        if (want_halt) {
          if (r != R_HALT) {
            fail = 1;
            if (nfail < 4) sf(&fl, "%s(%d, %d, 'expected halt, got %s')", nfail ? ", " : "", c, m, res_name[r]);
          }
        } else if (r != R_STOP) {
          fail = 1;
          if (nfail < 4)
            sf(&fl, "%s(%d, %d, 'expected stage 2m, got %s at %lld')", nfail ? ", " : "", c, m, res_name[r], steps);
        } else if (jw < ncnt && got.s && strcmp(got.s, cnt[jw].ctr)) {
          fail = 1;
          if (nfail < 4)
            sf(&fl, "%s(%d, %d, 'counter %s != c=%d %s')", nfail ? ", " : "", c, m, got.s, want_c, cnt[jw].ctr);
        } else if (jw == ncnt && got.s) {
          for (int a = 0; a < ncnt; a++)
            if (!strcmp(got.s, cnt[a].ctr)) fail = 1;
          if (fail && nfail < 4) sf(&fl, "%s(%d, %d, 'counter for c=3 equals an earlier one')", nfail ? ", " : "", c, m);
        }
        nfail += fail;
        free(got.s);
        free(ct.b);
      }
      if (rc) break;
    }
    if (!rc) {
      sf(o, "%s stages ok to n=%d (step %lld); counters {", nfail ? "FAIL" : "PASS", CK_NMAX, st_t[CK_NMAX]);
      for (int a = 0; a < ncnt; a++) sf(o, "%s%d: %s", a ? ", " : "", cnt[a].c, cnt[a].ctr);
      sf(o, "}; crafted fails %d: [%s]", nfail, fl.err ? "" : fl.s);
    }
    free(fl.s);
  }
done:
  for (int n = 0; n <= CK_NMAX; n++) free(ctrs[n]);
  free(t.b);
  return rc;
}

int e13_check_survivor(const char *line, char *out, int outlen) {
  if (!line || !out || outlen < 2) return E13_EINVAL;
  const char *tb = strstr(line, "table=");
  if (!tb) return E13_EINVAL;
  /* the line as given, without surrounding white space (check.py:168-169) */
  const char *a = line, *b = line + strlen(line);
  while (*a == ' ' || *a == '\t' || *a == '\n' || *a == '\r') a++;
  while (b > a && (b[-1] == ' ' || b[-1] == '\t' || b[-1] == '\n' || b[-1] == '\r')) b--;
  ck_cell rows[32][2];
  memset(rows, 0, sizeof rows);
  int unread[64], nun = 0, nq = 0;
  { /* parse (check.py:23-34): rows split at '|', cells at white space */
    const char *p = tb + 6;
    while (nq < 32) {
      for (int c = 0; c < 2; c++) {
        while (*p == ' ') p++;
        if (*p == '?') {
          unread[nun++] = 2 * nq + c;
          p++;
        }
        if (*p == '-') {
          rows[nq][c].halt = 1;
          p += 3;
        } else {
          rows[nq][c].w = p[0] - '0';
          rows[nq][c].d = p[1] == 'R' ? 1 : -1;
          char *e;
          rows[nq][c].nx = (int)strtol(p + 2, &e, 10);
          p = e;
        }
      }
      nq++;
      while (*p == ' ') p++;
      if (*p != '|') break;
      p++;
    }
  }
  int halts = 0;
  for (int i = 0; i < 2 * nq; i++) halts |= rows[i / 2][i % 2].halt;
  ck_str o = {(char *)malloc(4096), 0, 4096, 0};
  if (!o.s) return E13_ENOMEM;
  sf(&o, "%.*s\n", (int)(b - a), a);
  int rc = 0;
  const int nplace = halts ? 1 : nun;
  for (int k = 0; k < nplace && !rc; k++) {
    ck_cell r2[32][2];
    memcpy(r2, rows, sizeof rows);
    if (halts) sf(&o, "   halt@None: ");
    else {
      r2[unread[k] / 2][unread[k] % 2].halt = 1;
      sf(&o, "   halt@%d: ", unread[k]);
    }
    rc = ck_check((const ck_cell(*)[2])r2, nq, &o);
    sf(&o, "\n");
  }
  int len = (int)o.n;
  if (rc || o.err) len = E13_ENOMEM;
  else if ((int)o.n >= outlen) len = E13_EINVAL;
  else memcpy(out, o.s, o.n + 1);
  free(o.s);
  return len;
}
