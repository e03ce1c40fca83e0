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
      h.surv_bytes != sizeof(E13SurvRec) || memcmp(&h.prob, p, sizeof *p) != 0 || h.kept > h.nsurv)
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
