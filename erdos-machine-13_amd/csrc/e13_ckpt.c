/* e13_ckpt.c -- checkpoints of the host-loop search (the GPU driver's
 * checkpoint, clean stop and resume, src/e13gpu.py:19-23, 349-409).
 *
 * Between two launches the whole search state is the node stack [0, top), the
 * counters and the survivor records, so a checkpoint taken there is exact.
 * Two slots, <path>.a and <path>.b, are written in turn; the small pointer file
 * <path> names the last complete one and is replaced by rename only after the
 * slot is on disk, so a crash during a save leaves the previous checkpoint.
 * A slot file is a header (the problem it belongs to, stack depth, counters),
 * then the survivor records, then the stack nodes.
 */
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "e13_ctx.h"

#define CKPT_MAGIC 0x3154504B43333145ull /* "E13CKPT1" */
#define CKPT_CHUNK ((size_t)1 << 26)     /* bytes staged per copy */

typedef struct {
  uint64_t magic;
  uint32_t node_bytes, surv_bytes;
  e13_problem prob; /* the search the file belongs to */
  uint64_t top, peak, nsurv, kept;
  int64_t launches;
  double wall_ms;
  E13Counters ctr;
} ckpt_head;

int e13_set_checkpoint(e13_ctx *c, const char *path, double every_sec, int64_t max_launches) {
  if (!c) return E13_EINVAL;
  if (!path || !*path) {
    c->ckpt_path[0] = 0;
    return E13_OK;
  }
  if (strlen(path) + 8 > sizeof c->ckpt_path) return E13_EINVAL;
  snprintf(c->ckpt_path, sizeof c->ckpt_path, "%s", path);
  c->ckpt_every_s = every_sec > 0 ? every_sec : 1800.0; /* E13_CKPT_SEC default (e13gpu.py:351) */
  c->ckpt_max_launches = max_launches > 0 ? max_launches : 0;
  c->ckpt_slot = 0;
  return E13_OK;
}

/* Record counts above this mark a damaged header (and keep the size
   arithmetic below far from overflow). */
#define CKPT_MAX_RECORDS (1ull << 32)

static int write_all(FILE *f, const void *p, size_t n) { return fwrite(p, 1, n, f) == n ? 0 : -1; }

/* Device bytes [src, src+n) appended to f through a host staging buffer. */
static int dev_to_file(e13_ctx *c, FILE *f, const void *src, size_t n, char *stage) {
  for (size_t o = 0; o < n; o += CKPT_CHUNK) {
    const size_t k = n - o < CKPT_CHUNK ? n - o : CKPT_CHUNK;
    if (hipMemcpy(stage, (const char *)src + o, k, hipMemcpyDeviceToHost) != hipSuccess) return E13_EHIP;
    if (write_all(f, stage, k)) return E13_ECKPT;
  }
  return E13_OK;
}

static int file_to_dev(e13_ctx *c, FILE *f, void *dst, size_t n, char *stage) {
  for (size_t o = 0; o < n; o += CKPT_CHUNK) {
    const size_t k = n - o < CKPT_CHUNK ? n - o : CKPT_CHUNK;
    if (fread(stage, 1, k, f) != k) return E13_ECKPT;
    if (hipMemcpy((char *)dst + o, stage, k, hipMemcpyHostToDevice) != hipSuccess) return E13_EHIP;
  }
  return E13_OK;
}

int e13i_ckpt_save(e13_ctx *c, const e13_problem *p, size_t top, size_t peak, int64_t launches, double wall_ms) {
  char fn[1100], tmp[1100];
  const char sl = c->ckpt_slot ? 'b' : 'a';
  snprintf(fn, sizeof fn, "%s.%c", c->ckpt_path, sl);
  snprintf(tmp, sizeof tmp, "%s.tmp", c->ckpt_path);
  ckpt_head h;
  memset(&h, 0, sizeof h);
  h.magic = CKPT_MAGIC;
  h.node_bytes = (uint32_t)sizeof(E13Node);
  h.surv_bytes = (uint32_t)sizeof(E13SurvRec);
  h.prob = *p;
  h.top = top;
  h.peak = peak;
  h.launches = launches;
  h.wall_ms = wall_ms;
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(&h.ctr, c->d_ctr, sizeof h.ctr, hipMemcpyDeviceToHost));
  h.nsurv = h.ctr.nsurv;
  h.kept = h.nsurv < c->surv_cap ? h.nsurv : c->surv_cap;
  char *stage = (char *)malloc(CKPT_CHUNK);
  if (!stage) return E13_ENOMEM;
  FILE *f = fopen(fn, "wb");
  int rc = f ? E13_OK : E13_ECKPT;
  if (!rc && write_all(f, &h, sizeof h)) rc = E13_ECKPT;
  if (!rc) rc = dev_to_file(c, f, c->d_surv, h.kept * sizeof(E13SurvRec), stage);
  if (!rc) rc = dev_to_file(c, f, c->d_stack, top * sizeof(E13Node), stage);
  if (f) {
    if (!rc && (fflush(f) || fsync(fileno(f)))) rc = E13_ECKPT;
    if (fclose(f) && !rc) rc = E13_ECKPT;
  }
  free(stage);
  if (rc) return rc;
  /* then point to it: a crash before the rename keeps the old slot */
  FILE *pf = fopen(tmp, "w");
  if (!pf) return E13_ECKPT;
  fputc(sl, pf);
  fputc('\n', pf);
  if (fflush(pf) || fsync(fileno(pf))) rc = E13_ECKPT;
  if (fclose(pf) && !rc) rc = E13_ECKPT;
  if (!rc && rename(tmp, c->ckpt_path)) rc = E13_ECKPT;
  if (!rc) c->ckpt_slot ^= 1;
  return rc;
}

/* Resumes from the checkpoint named by the pointer file, if there is one:
   returns 1 with the stack, counters and survivor records on the device, 0
   when there is no checkpoint, or a negative code (E13_ECKPT: unreadable,
   damaged or written for another search). */
int e13i_ckpt_load(e13_ctx *c, const e13_problem *p, size_t *top, size_t *peak, int64_t *launches,
                   double *wall_ms) {
  char fn[1100], sl[8] = {0};
  HIPCHK(hipStreamSynchronize(c->stream)); /* the root and zeroed counters are queued there */
  FILE *pf = fopen(c->ckpt_path, "r");
  if (!pf) return errno == ENOENT ? 0 : E13_ECKPT;
  const int got = fgets(sl, sizeof sl, pf) != NULL;
  fclose(pf);
  if (!got || (sl[0] != 'a' && sl[0] != 'b')) return E13_ECKPT;
  snprintf(fn, sizeof fn, "%s.%c", c->ckpt_path, sl[0]);
  FILE *f = fopen(fn, "rb");
  if (!f) return E13_ECKPT;
  ckpt_head h;
  int rc = E13_OK;
  if (fread(&h, 1, sizeof h, f) != sizeof h || h.magic != CKPT_MAGIC || h.node_bytes != sizeof(E13Node) ||
      h.surv_bytes != sizeof(E13SurvRec) || memcmp(&h.prob, p, sizeof *p) != 0 || h.kept > h.nsurv ||
      h.kept > CKPT_MAX_RECORDS || h.top > CKPT_MAX_RECORDS)
    rc = E13_ECKPT;
  if (!rc) { /* the file holds exactly what the header says */
    const long long want = (long long)(sizeof h + h.kept * sizeof(E13SurvRec) + h.top * sizeof(E13Node));
    if (fseeko(f, 0, SEEK_END) || ftello(f) != want || fseeko(f, (off_t)sizeof h, SEEK_SET)) rc = E13_ECKPT;
  }
  if (!rc && h.top > c->stack_cap) rc = E13_ESTACK;
  if (!rc) rc = e13i_ensure_buf((void **)&c->d_surv, &c->surv_cap, h.kept > c->surv_cap ? h.kept : c->surv_cap,
                                sizeof(E13SurvRec));
  char *stage = rc ? NULL : (char *)malloc(CKPT_CHUNK);
  if (!rc && !stage) rc = E13_ENOMEM;
  if (!rc) rc = file_to_dev(c, f, c->d_surv, h.kept * sizeof(E13SurvRec), stage);
  if (!rc) rc = file_to_dev(c, f, c->d_stack, h.top * sizeof(E13Node), stage);
  free(stage);
  fclose(f);
  if (rc) return rc;
  h.ctr.item = h.ctr.nout = 0;
  if (hipMemcpy(c->d_ctr, &h.ctr, sizeof h.ctr, hipMemcpyHostToDevice) != hipSuccess) return E13_EHIP;
  *top = (size_t)h.top;
  *peak = (size_t)h.peak;
  *launches = h.launches;
  *wall_ms = h.wall_ms;
  c->ckpt_slot = sl[0] == 'a'; /* the next save goes to the other slot */
  return 1;
}

/* ------------------------------------------------------------------------ */
/* The persistent form.  A launch stops cleanly when the host raises a flag  */
/* in pinned memory: every wave pauses its running machines as resumable     */
/* nodes on its own stack and leaves, recording its stack range and the tag  */
/* of the last delivery it took.  The nodes left on the stacks, plus any     */
/* delivery still sitting in a mailbox, are the whole remaining search: they */
/* seed the next launch, or go to the checkpoint.                            */
/* ------------------------------------------------------------------------ */
#define CKPT3_MAGIC 0x3354504B43333145ull /* "E13CKPT3" */

typedef struct {
  uint64_t magic;
  uint32_t node_bytes, surv_bytes;
  int32_t np, pad;
  e13_problem prob[E13_MAXPROB];
  e13_counts counts[E13_MAXPROB];
  uint64_t nsurv, kept, nseeds;
} ckpt3_head;

int e13i_ckpt_stop_requested(e13_ctx *c) {
  char sp[1100];
  snprintf(sp, sizeof sp, "%s.stop", c->ckpt_path);
  return access(sp, F_OK) == 0;
}

void e13i_ckpt_stop_clear(e13_ctx *c) {
  char sp[1100];
  snprintf(sp, sizeof sp, "%s.stop", c->ckpt_path);
  remove(sp);
}

int e13i_ckpt3_prepare(e13_ctx *c, int total_waves) {
  if (!c->h_stop) {
    if (hipHostMalloc((void **)&c->h_stop, 64, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
      return E13_ENOMEM;
    if (hipHostGetDevicePointer((void **)&c->d_stop, c->h_stop, 0) != hipSuccess) return E13_EHIP;
  }
  *c->h_stop = 0u;
  return e13i_ensure_buf((void **)&c->d_wave_state, &c->wave_state_cap, 3 * (size_t)total_waves, sizeof(int64_t));
}

/* The nodes a stopped launch left, gathered into d_seeds (count in *nseeds):
   every wave's stack range and the deliveries still in mailboxes (*n_left of
   them), then n_extra nodes from extra (the narrow form's wide list). */
int e13i_ckpt3_gather(e13_ctx *c, int total_waves, int stk_cap, const E13Node *extra, size_t n_extra, size_t *n_left,
                      size_t *nseeds) {
  const size_t W = (size_t)total_waves;
  int64_t *ws = (int64_t *)malloc(3 * W * sizeof(int64_t));
  E13Mail *mail = (E13Mail *)malloc(W * sizeof(E13Mail));
  int rc = ws && mail ? E13_OK : E13_ENOMEM;
  if (!rc && (hipMemcpy(ws, c->d_wave_state, 3 * W * sizeof(int64_t), hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(mail, c->d_mail, W * sizeof(E13Mail), hipMemcpyDeviceToHost) != hipSuccess))
    rc = E13_EHIP;
  size_t total = 0, deliveries = 0;
  for (size_t w = 0; !rc && w < W; w++) {
    const int64_t bot = ws[3 * w], top = ws[3 * w + 1];
    if (top < bot || top - bot > stk_cap) rc = E13_ESTACK;
    total += (size_t)(top - bot);
    const uint64_t f = mail[w].flag;
    if (f != 0 && f != E13_MAIL_DONE && f != (uint64_t)ws[3 * w + 2]) deliveries++; /* not taken */
  }
  if (!rc) rc = e13i_ensure_buf((void **)&c->d_seeds, &c->seeds_cap, total + deliveries + n_extra + 1, sizeof(E13Node));
  size_t off = 0;
  for (size_t w = 0; !rc && w < W; w++) {
    const int64_t bot = ws[3 * w], top = ws[3 * w + 1];
    const E13Node *stack = c->d_stacks + w * (size_t)stk_cap;
    for (int64_t i = bot; i < top;) { /* the stack is a ring of stk_cap nodes */
      const int64_t r = i % stk_cap, n = (top - i) < (stk_cap - r) ? (top - i) : (stk_cap - r);
      if (hipMemcpyAsync(c->d_seeds + off, stack + r, (size_t)n * sizeof(E13Node), hipMemcpyDeviceToDevice,
                         c->stream) != hipSuccess)
        rc = E13_EHIP;
      off += (size_t)n;
      i += n;
    }
  }
  for (size_t w = 0; !rc && w < W; w++) {
    const uint64_t f = mail[w].flag;
    if (f == 0 || f == E13_MAIL_DONE || f == (uint64_t)ws[3 * w + 2]) continue;
    E13Node nd;
    uint32_t *dw = (uint32_t *)&nd;
    for (int i = 0; i < E13_NODE_WORDS; i++) {
      if ((uint32_t)(mail[w].w[i] >> 32) != (uint32_t)f) rc = E13_EHIP; /* every delivery is complete by now */
      dw[i] = (uint32_t)mail[w].w[i];
    }
    if (!rc && (hipMemcpy(c->d_seeds + off, &nd, sizeof nd, hipMemcpyHostToDevice) != hipSuccess)) rc = E13_EHIP;
    off++;
  }
  *n_left = off;
  if (!rc && n_extra &&
      hipMemcpyAsync(c->d_seeds + off, extra, n_extra * sizeof(E13Node), hipMemcpyDeviceToDevice, c->stream) !=
          hipSuccess)
    rc = E13_EHIP;
  off += n_extra;
  if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = E13_EHIP;
  free(ws);
  free(mail);
  *nseeds = off;
  return rc;
}

int e13i_ckpt3_save(e13_ctx *c, const e13_problem *ps, int np, const e13_counts *out, size_t nseeds) {
  char fn[1100], tmp[1100];
  const char sl = c->ckpt_slot ? 'b' : 'a';
  snprintf(fn, sizeof fn, "%s.%c", c->ckpt_path, sl);
  snprintf(tmp, sizeof tmp, "%s.tmp", c->ckpt_path);
  ckpt3_head h;
  memset(&h, 0, sizeof h);
  h.magic = CKPT3_MAGIC;
  h.node_bytes = (uint32_t)sizeof(E13Node);
  h.surv_bytes = (uint32_t)sizeof(E13SurvRec);
  h.np = np;
  memcpy(h.prob, ps, (size_t)np * sizeof *ps);
  memcpy(h.counts, out, (size_t)np * sizeof *out);
  E13Counters ctr;
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(&ctr, c->d_ctr, sizeof ctr, hipMemcpyDeviceToHost));
  h.nsurv = ctr.nsurv;
  h.kept = h.nsurv < c->surv_cap ? h.nsurv : c->surv_cap;
  h.nseeds = nseeds;
  char *stage = (char *)malloc(CKPT_CHUNK);
  if (!stage) return E13_ENOMEM;
  FILE *f = fopen(fn, "wb");
  int rc = f ? E13_OK : E13_ECKPT;
  if (!rc && write_all(f, &h, sizeof h)) rc = E13_ECKPT;
  if (!rc) rc = dev_to_file(c, f, c->d_surv, h.kept * sizeof(E13SurvRec), stage);
  if (!rc) rc = dev_to_file(c, f, c->d_seeds, nseeds * sizeof(E13Node), stage);
  if (f) {
    if (!rc && (fflush(f) || fsync(fileno(f)))) rc = E13_ECKPT;
    if (fclose(f) && !rc) rc = E13_ECKPT;
  }
  free(stage);
  if (rc) return rc;
  FILE *pf = fopen(tmp, "w");
  if (!pf) return E13_ECKPT;
  fputc(sl, pf);
  fputc('\n', pf);
  if (fflush(pf) || fsync(fileno(pf))) rc = E13_ECKPT;
  if (fclose(pf) && !rc) rc = E13_ECKPT;
  if (!rc && rename(tmp, c->ckpt_path)) rc = E13_ECKPT;
  if (!rc) c->ckpt_slot ^= 1;
  return rc;
}

/* Resume: 1 with the counts in out[], the survivor records and the nodes in
   d_seeds; 0 when there is no checkpoint; E13_ECKPT for another search's or a
   damaged file (or one written by the host-loop form). */
int e13i_ckpt3_load(e13_ctx *c, const e13_problem *ps, int np, e13_counts *out, size_t *nseeds) {
  char fn[1100], sl[8] = {0};
  HIPCHK(hipStreamSynchronize(c->stream));
  FILE *pf = fopen(c->ckpt_path, "r");
  if (!pf) return errno == ENOENT ? 0 : E13_ECKPT;
  const int got = fgets(sl, sizeof sl, pf) != NULL;
  fclose(pf);
  if (!got || (sl[0] != 'a' && sl[0] != 'b')) return E13_ECKPT;
  snprintf(fn, sizeof fn, "%s.%c", c->ckpt_path, sl[0]);
  FILE *f = fopen(fn, "rb");
  if (!f) return E13_ECKPT;
  ckpt3_head h;
  int rc = E13_OK;
  if (fread(&h, 1, sizeof h, f) != sizeof h || h.magic != CKPT3_MAGIC || h.node_bytes != sizeof(E13Node) ||
      h.surv_bytes != sizeof(E13SurvRec) || h.np != np || memcmp(h.prob, ps, (size_t)np * sizeof *ps) != 0 ||
      h.kept > h.nsurv || h.kept > CKPT_MAX_RECORDS || h.nseeds > CKPT_MAX_RECORDS)
    rc = E13_ECKPT;
  if (!rc) {
    const long long want = (long long)(sizeof h + h.kept * sizeof(E13SurvRec) + h.nseeds * sizeof(E13Node));
    if (fseeko(f, 0, SEEK_END) || ftello(f) != want || fseeko(f, (off_t)sizeof h, SEEK_SET)) rc = E13_ECKPT;
  }
  if (!rc) rc = e13i_ensure_buf((void **)&c->d_surv, &c->surv_cap, h.kept > c->surv_cap ? h.kept : c->surv_cap,
                                sizeof(E13SurvRec));
  if (!rc) rc = e13i_ensure_buf((void **)&c->d_seeds, &c->seeds_cap, h.nseeds + 1, sizeof(E13Node));
  char *stage = rc ? NULL : (char *)malloc(CKPT_CHUNK);
  if (!rc && !stage) rc = E13_ENOMEM;
  if (!rc) rc = file_to_dev(c, f, c->d_surv, h.kept * sizeof(E13SurvRec), stage);
  if (!rc) rc = file_to_dev(c, f, c->d_seeds, h.nseeds * sizeof(E13Node), stage);
  free(stage);
  fclose(f);
  if (rc) return rc;
  E13Counters ctr;
  memset(&ctr, 0, sizeof ctr);
  ctr.nsurv = (uint32_t)h.nsurv;
  if (hipMemcpy(c->d_ctr, &ctr, sizeof ctr, hipMemcpyHostToDevice) != hipSuccess) return E13_EHIP;
  memcpy(out, h.counts, (size_t)np * sizeof *out);
  *nseeds = (size_t)h.nseeds;
  c->ckpt_slot = sl[0] == 'a';
  return 1;
}
