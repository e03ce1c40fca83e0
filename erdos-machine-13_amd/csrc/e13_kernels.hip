// e13_kernels.hip -- gfx950 kernels for the TNF enumeration + simulate-and-filter
// hot path (reference: src/e13search.c run()/explore(), src/e13gpu.py expand()).
//
// One machine per lane of a 64-wide wavefront.
//
// LDS, per wave:
//   rows  16 x 64 words at a 4096-byte aligned base: the word of state s for
//         lane l holds the two rule entries (symbol 0 in bits 0..15, symbol 1 in
//         bits 16..31).  An entry is  write | dir << 1 | next << 8, or 0x0F00 |
//         kind << 12 for a cell that cannot be stepped (kind 1 halt, 2 unread
//         editable cell, 3 free cell not chosen yet).  Because the row base has
//         no bits in 0x0F00, the LDS address of the next row is
//         (entry & 0x0F00) | lane_base: one VALU op.
//   tape  43 x 64 words: word w of lane l at w*256 + l*4 (every lane in its own
//         bank, so tape reads and writes never conflict).
//
// Per step (the reference's loop body, e13search.c:136-175, same order):
//   stage gate -> deadline/cap -> read -> unbound/halt stop -> write -> move ->
//   window -> next state, t++.
// The stage gate is an incremental count of the ones at cells >= -2 compared
// with 2^nt, with the stage-1 state and head as a prefilter and an exact block
// check behind it (GATE_COUNT), or the reference's 64-bit Zobrist block hash
// (GATE_HASH, the verification build).  The counter hash hc is rebuilt from
// the counter cells at an accepted stage start, so ch[] is the reference's.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "e13_dev.h"

#define WAVE 64
#define ROWS_BYTES 4096
#define TAPE_BYTES (E13_TW * 256)
#define KEYB_WORDS 1376

enum { EV_NONE = 0, EV_DEAD = 1, EV_SURV = 2, EV_STOP = 3, EV_CAP = 4 };
enum { GATE_COUNT = 0, GATE_HASH = 1 };

namespace {

__device__ __forceinline__ uint32_t lds_ld(const char *sb, int off) { return *(const uint32_t *)(sb + off); }
__device__ __forceinline__ void lds_st(char *sb, int off, uint32_t v) { *(uint32_t *)(sb + off) = v; }

__device__ __forceinline__ uint32_t enc_entry(int c) {
  if (c < 60) return (uint32_t)(c & 3) | ((uint32_t)(c >> 2) << 8);
  uint32_t kind = c == E13_C_HALT ? 1u : (c == E13_C_EDIT ? 2u : 3u);
  return 0x0F00u | (kind << 12);
}

__device__ __forceinline__ int dec_entry(uint32_t e) {
  e &= 0xFFFFu;
  if ((e & 0x0F00u) == 0x0F00u) {
    uint32_t kind = (e >> 12) & 3u;
    return kind == 1 ? E13_C_HALT : (kind == 2 ? E13_C_EDIT : E13_C_FREE);
  }
  return (int)((((e >> 8) & 15u) << 2) | (e & 3u));
}

__device__ __forceinline__ int cval(int n) { return (n >= 3) + (n >= 9); } // e13search.c:71

struct Lane {
  int q, row, t, tstop, deadline, tcap, dlcap;
  int ones, target, row0, qA, qR;
  int nt, stage, st0, q0, flags;
  uint64_t hb, tgt;
  uint64_t ch0, ch1, ch2;
};

struct Ctx {
  char *sb;
  int rbl, tbl;       // row and tape byte bases of this lane
  int keyb;           // byte base of the per-workgroup block keys (hash gate)
  const uint64_t *key;
  const uint64_t *tgt;
  const int32_t *dl;
};

__device__ __forceinline__ int lane_deadline(const Lane &L, const Ctx &c, int n) {
  return L.dlcap >= 0 ? L.dlcap : c.dl[n];
}

// Gate registers for target stage nt (after a load or an accepted stage).
__device__ __forceinline__ void arm_stage(Lane &L, const Ctx &c) {
  L.target = 1 << L.nt;
  L.row0 = c.rbl + L.st0 * 256;
  L.qA = (L.flags & 1) ? L.q0 : -1000;
  L.qR = (L.flags & 2) ? L.q0 - 2 + (1 << L.nt) : -1000;
  L.deadline = lane_deadline(L, c, L.nt);
  L.tstop = min(L.deadline + 1, L.tcap);
  L.tgt = c.tgt[L.nt];
}

// Cells -1 .. 2^n-2 (q 62 .. 2^n+61) all ones.  With exactly 2^n ones at
// q >= 61 this is the reference's block equality hb == TGT[n].
__device__ __forceinline__ bool block_exact(const Ctx &c, int n) {
  const int qlo = 62, qhi = (1 << n) + 61;
  for (int w = qlo >> 5; w <= (qhi >> 5); ++w) {
    uint32_t m = 0xFFFFFFFFu;
    if (w == (qlo >> 5)) m &= 0xFFFFFFFFu << (qlo & 31);
    if (w == (qhi >> 5)) m &= 0xFFFFFFFFu >> (31 - (qhi & 31));
    if ((lds_ld(c.sb, c.tbl + w * 256) & m) != m) return false;
  }
  return true;
}

// hc of the reference: XOR of KEY[cell+64] over the set cells -63..-3
// (q 0..60, KEY index q+1), e13search.c:170.
__device__ __forceinline__ uint64_t counter_hash(const Ctx &c) {
  uint32_t w0 = lds_ld(c.sb, c.tbl);
  uint32_t w1 = lds_ld(c.sb, c.tbl + 256) & 0x1FFFFFFFu;
  uint64_t h = 0;
  while (w0) {
    int b = __builtin_ctz(w0);
    h ^= c.key[b + 1];
    w0 &= w0 - 1;
  }
  while (w1) {
    int b = __builtin_ctz(w1);
    h ^= c.key[b + 33];
    w1 &= w1 - 1;
  }
  return h;
}

// The stage-start rule (e13search.c:136-161) at a gate hit.
template <int GATE>
__device__ __forceinline__ int stage_event(Lane &L, const Ctx &c) {
  const int n = L.nt;
  if (GATE == GATE_COUNT) {
    if (!block_exact(c, n)) return EV_NONE;
  }
  const int st = (L.row - c.rbl) >> 8;
  int a_ok = L.flags & 1, r_ok = (L.flags >> 1) & 1;
  if (n > 1) {
    if (st != L.st0) return EV_NONE;
    if (L.q != L.q0) a_ok = 0;
    if (L.q - (1 << n) != L.q0 - 2) r_ok = 0;
    if (!a_ok && !r_ok) return EV_NONE;
  }
  const uint64_t hc = counter_hash(c);
  const int cc = cval(n);
  // value selects (a conditional over the fields would select addresses and
  // keep the lane state out of registers)
  uint64_t chc = L.ch0;
  if (cc == 1) chc = L.ch1;
  if (cc == 2) chc = L.ch2;
  if (((L.flags >> (2 + cc)) & 1) && chc != hc) return EV_DEAD;
  if ((cc != 0 && (L.flags & 4) && L.ch0 == hc) || (cc != 1 && (L.flags & 8) && L.ch1 == hc) ||
      (cc != 2 && (L.flags & 16) && L.ch2 == hc))
    return EV_DEAD;
  L.flags |= 4 << cc;
  if (cc == 0) L.ch0 = hc;
  else if (cc == 1) L.ch1 = hc;
  else L.ch2 = hc;
  if (n == 1) {
    L.st0 = st;
    L.q0 = L.q;
  }
  L.flags = (L.flags & ~3) | a_ok | (r_ok << 1);
  L.stage = n;
  if (n == E13_NS) return EV_SURV;
  L.nt = n + 1;
  arm_stage(L, c);
  return EV_NONE;
}

// Up to `slice` steps.  Returns an event; for EV_STOP the cell is in evcell.
// Within a round every running lane takes one step per iteration, so the step
// count is the round's start count plus the (wave-uniform) iteration index:
// t needs no per-lane increment, and the deadline/cap test compares a
// per-lane budget with that index.
template <int GATE>
__device__ __forceinline__ int run_slice(Lane &L, const Ctx &c, int slice, int &evcell) {
  const int t_round = L.t;
  int left = L.tstop - t_round; // steps this lane may still take before the deadline/cap test fires
  int ev = EV_NONE;
  int r = 0;
  for (; r < slice; ++r) {
    // stage gate (e13search.c:136)
    const bool hit = GATE == GATE_COUNT ? L.ones == L.target : L.hb == L.tgt;
    if (hit) {
      const bool pre =
          GATE == GATE_HASH || ((L.nt == 1) | ((L.row == L.row0) & ((L.q == L.qA) | (L.q == L.qR))));
      if (pre) {
        L.t = t_round + r;
        ev = stage_event<GATE>(L, c);
        if (ev) break;
        left = L.tstop - t_round;
      }
    }
    const uint32_t qb = (uint32_t)L.q & 31u;
    const int wa = (int)(__builtin_amdgcn_ubfe((uint32_t)L.q, 5u, 6u) << 8) + c.tbl;
    uint32_t word = lds_ld(c.sb, wa);
    const uint32_t pair = lds_ld(c.sb, L.row);
    const uint32_t sym = __builtin_amdgcn_ubfe(word, qb, 1u);
    const uint32_t e = pair >> (sym << 4);
    const uint32_t nx = e & 0x0F00u;
    const bool timeup = r >= left;
    if (timeup | (nx == 0x0F00u)) {
      if (timeup) ev = t_round + r > L.deadline ? EV_DEAD : EV_CAP; // deadline (e13search.c:162) or launch cap
      else if (((e >> 12) & 3u) == 1u) ev = EV_DEAD;                 // halt read (e13search.c:166)
      else {                                                         // first read of an unbound cell (:164)
        evcell = ((L.row - c.rbl) >> 7) + (int)sym;
        ev = EV_STOP;
      }
      break;
    }
    const uint32_t w = e & 1u;
    word ^= (w ^ sym) << qb;
    lds_st(c.sb, wa, word);
    if (GATE == GATE_COUNT) {
      L.ones += L.q >= E13_QBLK ? (int)w - (int)sym : 0;
    } else {
      const uint64_t kb = *(const uint64_t *)(c.sb + c.keyb + L.q * 8);
      L.hb ^= kb & (0ull - (uint64_t)(w ^ sym));
    }
    L.q += (int)(e & 2u) - 1;
    if ((unsigned)L.q > (unsigned)E13_QMAX) { // off the window; the step is not counted (e13search.c:173)
      ev = EV_DEAD;
      break;
    }
    L.row = (int)(nx | (uint32_t)c.rbl);
  }
  L.t = t_round + r;
  return ev;
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ uint32_t rank_in(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Per-workgroup block keys for the hash gate: keyb[q] = KEY[q+1] for q >= 61, else 0.
__device__ __forceinline__ void fill_keyb(char *sb, int keyb, const uint64_t *key) {
  for (int q = threadIdx.x; q < KEYB_WORDS; q += blockDim.x) {
    uint64_t k = (q >= E13_QBLK && q + 1 < 1364) ? key[q + 1] : 0ull;
    *(uint64_t *)(sb + keyb + q * 8) = k;
  }
}

__device__ __forceinline__ int sym_next(int code, int nintro, int sym_lo, int sym_hi) { // e13search.c:206-212
  if (sym_lo < 0) return nintro;
  const int qq = code >> 2;
  if (qq < sym_lo || qq > sym_hi) return nintro;
  if (qq - sym_lo < nintro) return nintro;
  return qq - sym_lo == nintro ? nintro + 1 : -1;
}

// The lane's rows from 32 codes; `pend` gets `val` (if pend >= 0); with no
// budget left, unread editable cells take their base values.
__device__ __forceinline__ void build_rows(const Ctx &c, const uint8_t *tab, int nq, int pend, int val, int budget,
                                           const int8_t *base) {
  for (int r = 0; r < nq; ++r) {
    int c0 = tab[2 * r], c1 = tab[2 * r + 1];
    if (2 * r == pend) c0 = val;
    if (2 * r + 1 == pend) c1 = val;
    if (budget == 0) {
      if (c0 == E13_C_EDIT) c0 = base[2 * r] < 0 ? E13_C_HALT : base[2 * r];
      if (c1 == E13_C_EDIT) c1 = base[2 * r + 1] < 0 ? E13_C_HALT : base[2 * r + 1];
    }
    lds_st(c.sb, c.rbl + r * 256, enc_entry(c0) | (enc_entry(c1) << 16));
  }
}

// Codes of cells 4i..4i+3 packed little-endian (rows past nq read as free).
__device__ __forceinline__ uint32_t codes_word(const Ctx &c, int nq, int i) {
  uint32_t v = 0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int r = 2 * i + h;
    uint32_t lo = E13_C_FREE, hi = E13_C_FREE;
    if (r < nq) {
      const uint32_t pr = lds_ld(c.sb, c.rbl + r * 256);
      lo = (uint32_t)dec_entry(pr);
      hi = (uint32_t)dec_entry(pr >> 16);
    }
    v |= (lo | (hi << 8)) << (16 * h);
  }
  return v;
}

} // namespace

// ---------------------------------------------------------------------------
// Expansion: each work item is (input node, slot).  Slot 0 is the "keep" child
// (the base value of an editable cell, e13search.c:238-241), slot 1+v binds the
// pending cell to move v (e13search.c:242-251).  A lane runs its child until it
// dies, survives, stops at the next unbound read, or hits the per-launch cap;
// live children go to `out`.  REFILL=1: persistent lanes take new items
// whenever they finish (wave-wide ballot, one atomic per wave); REFILL=0: one
// item per lane (the plain form).
// ---------------------------------------------------------------------------
template <int GATE, int REFILL>
__global__ void __launch_bounds__(256) k_expand(E13ExpandArgs a) {
  extern __shared__ __align__(16) uint32_t smem[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  Ctx c;
  c.sb = (char *)smem;
  c.rbl = wave * ROWS_BYTES + lane * 4;
  c.tbl = nwaves * ROWS_BYTES + wave * TAPE_BYTES + lane * 4;
  c.keyb = nwaves * (ROWS_BYTES + TAPE_BYTES);
  c.key = a.key;
  c.tgt = a.tgt;
  c.dl = a.pb->dl;
  // per-wave stage histogram: 16 words after the tapes and block keys
  uint32_t *whist = (uint32_t *)(c.sb + c.keyb + (GATE == GATE_HASH ? KEYB_WORDS * 8 : 0)) + wave * 16;
  if (lane < 16) whist[lane] = 0;
  if (GATE == GATE_HASH) fill_keyb(c.sb, c.keyb, a.key);
  __syncthreads();

  const uint32_t nitems = a.nin * a.nslot;
  Lane L;
  L.dlcap = -1;
  int active = 0, t0 = 0, budget = 0, nintro = 0, front = 0, depth = 0, nbound = 0;
  uint64_t my_steps = 0;
  uint32_t my_leaves = 0, my_surv = 0;
  uint64_t w_slots = 0, w_rounds = 0;
  bool exhausted = false;
  bool first = true;

  for (;;) {
    // ---- refill ----
    if (REFILL || first) {
      uint64_t need = ballot(!active);
      for (int tries = 0; need != 0 && !exhausted && tries < a.refill_tries; ++tries) {
        uint32_t k;
        if (REFILL) {
          const int cnt = __popcll(need);
          const int leader = __ffsll((long long)need) - 1;
          uint32_t b0 = 0;
          if (lane == leader) b0 = atomicAdd(&a.ctr->item, (uint32_t)cnt);
          b0 = __shfl(b0, leader);
          if (b0 + (uint32_t)cnt >= nitems) exhausted = true;
          k = b0 + rank_in(need);
        } else {
          k = blockIdx.x * blockDim.x + threadIdx.x;
          exhausted = true;
        }
        if (!active && k < nitems) {
          const uint32_t ni = k / a.nslot, slot = k - ni * a.nslot;
          const E13Node *nd = a.in + ni;
          const int pend = nd->pend;
          int nb = nd->budget, nin2 = nd->nintro, val = -1, ok = 1;
          if (pend < 0) {
            ok = slot == 0;
          } else {
            const int orig = a.pb->base[pend];
            if (slot == 0) {
              if (orig == -2) ok = 0;
              else if (orig == -1) { // keeping a halt: dies at once, 0 steps (e13search.c:164-166)
                atomicAdd(&whist[nd->stage], 1u);
                my_leaves++;
                ok = 0;
              } else val = orig;
            } else {
              const int v = (int)slot - 1;
              if (!(orig == -2 || nb > 0) || v == orig) ok = 0;
              else {
                const int n2 = sym_next(v, nin2, a.pb->sym_lo, a.pb->sym_hi);
                if (n2 < 0) ok = 0;
                else {
                  if (orig != -2) nb -= 1;
                  nin2 = n2;
                  val = v;
                }
              }
            }
          }
          if (ok) {
            L.q = nd->q;
            L.row = c.rbl + nd->st * 256;
            L.t = nd->t;
            L.nt = nd->nt;
            L.stage = nd->stage;
            L.st0 = nd->st0;
            L.q0 = nd->q0;
            L.flags = nd->flags;
            L.ones = nd->ones;
            L.hb = nd->hb;
            L.ch0 = nd->ch[0];
            L.ch1 = nd->ch[1];
            L.ch2 = nd->ch[2];
            L.tcap = L.t + a.cap;
            arm_stage(L, c);
            t0 = L.t;
            budget = nb;
            nintro = nin2;
            front = nd->front;
            depth = nd->depth + (pend >= 0);
            nbound = nd->nbound + (pend >= 0);
            build_rows(c, nd->tab, a.nq, pend, val, budget, a.pb->base);
            const uint4 *src = (const uint4 *)nd->tape;
            for (int w4 = 0; w4 < E13_TW_PAD / 4; ++w4) {
              const uint4 v4 = src[w4];
              lds_st(c.sb, c.tbl + (4 * w4 + 0) * 256, v4.x);
              lds_st(c.sb, c.tbl + (4 * w4 + 1) * 256, v4.y);
              lds_st(c.sb, c.tbl + (4 * w4 + 2) * 256, v4.z);
              if (4 * w4 + 3 < E13_TW) lds_st(c.sb, c.tbl + (4 * w4 + 3) * 256, v4.w);
            }
            active = 1;
          }
        }
        need = ballot(!active);
      }
    }
    first = false;
    const uint64_t act = ballot(active);
    if (act == 0) {
      if (exhausted) break;
      continue;
    }
    w_rounds++;
    w_slots += (uint64_t)__popcll(act);
    // ---- a round of steps ----
    int ev = EV_NONE, evcell = -1;
    if (active) ev = run_slice<GATE>(L, c, a.slice, evcell);
    // ---- finished machines ----
    if (active && ev != EV_NONE) {
      active = 0;
      my_steps += (uint64_t)(L.t - t0);
      if (ev == EV_DEAD) {
        my_leaves++;
        atomicAdd(&whist[L.stage], 1u);
      } else if (ev == EV_SURV) {
        my_leaves++;
        my_surv++;
        const uint32_t s = atomicAdd(&a.ctr->nsurv, 1u);
        if (s < a.surv_cap) {
          E13SurvRec *o = a.surv + s;
          uint32_t *tw = (uint32_t *)o->tab;
#pragma unroll
          for (int i = 0; i < E13_MAXCELLS / 4; ++i) tw[i] = codes_word(c, a.nq, i);
          o->st0 = L.st0;
          o->head0 = L.q0 - E13_QOFF;
          o->abs_ok = L.flags & 1;
          o->rel_ok = (L.flags >> 1) & 1;
          o->steps = L.t;
          o->pad = 0;
          o->ch[0] = L.ch0;
          o->ch[1] = L.ch1;
          o->ch[2] = L.ch2;
        }
      } else { // EV_STOP or EV_CAP: a live child
        const uint32_t s = atomicAdd(&a.ctr->nout, 1u);
        if (s < a.out_cap) {
          E13Node *o = a.out + s;
          o->q = L.q;
          o->st = (L.row - c.rbl) >> 8;
          o->t = L.t;
          o->nt = L.nt;
          o->stage = L.stage;
          o->st0 = L.st0;
          o->q0 = L.q0;
          o->flags = L.flags;
          o->ones = L.ones;
          o->budget = budget;
          o->nintro = nintro;
          o->pend = ev == EV_STOP ? evcell : -1;
          o->front = front;
          o->depth = depth;
          o->nbound = nbound;
          o->pad = 0;
          o->hb = L.hb;
          o->ch[0] = L.ch0;
          o->ch[1] = L.ch1;
          o->ch[2] = L.ch2;
          uint32_t *tw = (uint32_t *)o->tab;
#pragma unroll
          for (int i = 0; i < E13_MAXCELLS / 4; ++i) tw[i] = codes_word(c, a.nq, i);
          uint4 *dst = (uint4 *)o->tape;
          for (int w4 = 0; w4 < E13_TW_PAD / 4; ++w4) {
            uint4 v4;
            v4.x = lds_ld(c.sb, c.tbl + (4 * w4 + 0) * 256);
            v4.y = lds_ld(c.sb, c.tbl + (4 * w4 + 1) * 256);
            v4.z = lds_ld(c.sb, c.tbl + (4 * w4 + 2) * 256);
            v4.w = (4 * w4 + 3 < E13_TW) ? lds_ld(c.sb, c.tbl + (4 * w4 + 3) * 256) : 0u;
            dst[w4] = v4;
          }
        } else {
          atomicOr(&a.ctr->overflow, 1u);
        }
      }
    }
    if (!REFILL && ballot(active) == 0) break;
  }
  // ---- flush counters ----
  const uint64_t s_steps = wave_sum64(my_steps);
  const uint64_t s_leaves = wave_sum64((uint64_t)my_leaves);
  const uint64_t s_surv = wave_sum64((uint64_t)my_surv);
  if (lane == 0) {
    atomicAdd((unsigned long long *)&a.ctr->steps, (unsigned long long)s_steps);
    atomicAdd((unsigned long long *)&a.ctr->leaves, (unsigned long long)s_leaves);
    if (s_surv) atomicAdd((unsigned long long *)&a.ctr->survivors, (unsigned long long)s_surv);
    atomicAdd((unsigned long long *)&a.ctr->rounds, (unsigned long long)w_rounds);
    atomicAdd((unsigned long long *)&a.ctr->lane_slots, (unsigned long long)w_slots);
  }
  __syncthreads();
  if (lane <= E13_NS && whist[lane])
    atomicAdd((unsigned long long *)&a.ctr->hist[lane], (unsigned long long)whist[lane]);
}

// ---------------------------------------------------------------------------
// Complete tables from the blank tape (run_full, e13search.c:428-433): the
// mutant sweep and the synthetic stress batch.  Persistent lanes with refill.
// ---------------------------------------------------------------------------
template <int GATE>
__global__ void __launch_bounds__(256) k_batch(E13BatchArgs a) {
  extern __shared__ __align__(16) uint32_t smem[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  Ctx c;
  c.sb = (char *)smem;
  c.rbl = wave * ROWS_BYTES + lane * 4;
  c.tbl = nwaves * ROWS_BYTES + wave * TAPE_BYTES + lane * 4;
  c.keyb = nwaves * (ROWS_BYTES + TAPE_BYTES);
  c.key = a.key;
  c.tgt = a.tgt;
  c.dl = a.dl;
  if (GATE == GATE_HASH) fill_keyb(c.sb, c.keyb, a.key);
  __syncthreads();
  Lane L;
  int active = 0;
  uint32_t k = 0;
  uint64_t my_steps = 0, w_slots = 0, w_rounds = 0;
  bool exhausted = false;
  for (;;) {
    const uint64_t need = ballot(!active);
    if (need != 0 && !exhausted) {
      const int cnt = __popcll(need);
      const int leader = __ffsll((long long)need) - 1;
      uint32_t b0 = 0;
      if (lane == leader) b0 = atomicAdd(&a.ctr->item, (uint32_t)cnt);
      b0 = __shfl(b0, leader);
      if (b0 + (uint32_t)cnt >= (uint32_t)a.n) exhausted = true;
      if (!active) {
        const uint32_t kk = b0 + rank_in(need);
        if (kk < (uint32_t)a.n) {
          k = kk;
          build_rows(c, a.tabs + (size_t)k * E13_MAXCELLS, a.nq, -1, -1, 1, nullptr);
          for (int w = 0; w < E13_TW; ++w) lds_st(c.sb, c.tbl + w * 256, 0u);
          L.q = E13_QOFF;
          L.row = c.rbl;
          L.t = 0;
          L.nt = 1;
          L.stage = 0;
          L.st0 = -1;
          L.q0 = E13_QOFF;
          L.flags = 3;
          L.ones = 0;
          L.hb = 0;
          L.ch0 = L.ch1 = L.ch2 = 0;
          L.dlcap = a.caps ? a.caps[k] : -1;
          L.tcap = 0x7FFFFFFF;
          arm_stage(L, c);
          active = 1;
        }
      }
    }
    const uint64_t act = ballot(active);
    if (act == 0) {
      if (exhausted) break;
      continue;
    }
    w_rounds++;
    w_slots += (uint64_t)__popcll(act);
    int ev = EV_NONE, evcell = -1;
    if (active) ev = run_slice<GATE>(L, c, a.slice, evcell);
    if (active && ev != EV_NONE) {
      active = 0;
      my_steps += (uint64_t)L.t;
      E13LeafRec *o = a.out + k;
      o->result = ev == EV_SURV ? -2 : -1;
      o->stage = L.stage;
      o->steps = L.t;
      o->st0 = L.st0;
      o->head0 = L.q0 - E13_QOFF;
      o->abs_ok = L.flags & 1;
      o->rel_ok = (L.flags >> 1) & 1;
      o->pad = ev;
      o->ch[0] = L.ch0;
      o->ch[1] = L.ch1;
      o->ch[2] = L.ch2;
    }
  }
  const uint64_t s_steps = wave_sum64(my_steps);
  if (lane == 0) {
    atomicAdd((unsigned long long *)&a.ctr->steps, (unsigned long long)s_steps);
    atomicAdd((unsigned long long *)&a.ctr->rounds, (unsigned long long)w_rounds);
    atomicAdd((unsigned long long *)&a.ctr->lane_slots, (unsigned long long)w_slots);
  }
}

// ---------------------------------------------------------------------------
// Thin C ABI over the launches.  Shapes are checked by the caller (host C).
// ---------------------------------------------------------------------------
extern "C" {

size_t e13k_lds_bytes(int wg_threads, int gate) {
  const int nw = wg_threads / WAVE;
  return (size_t)nw * (ROWS_BYTES + TAPE_BYTES + 64) + (gate == GATE_HASH ? KEYB_WORDS * 8 : 0);
}

int e13k_launch_expand(const E13ExpandArgs *a, int grid, int wg_threads, int gate, int refill, hipStream_t s) {
  const size_t lds = e13k_lds_bytes(wg_threads, gate);
  if (gate == GATE_HASH) {
    if (refill) hipLaunchKernelGGL((k_expand<GATE_HASH, 1>), dim3(grid), dim3(wg_threads), lds, s, *a);
    else hipLaunchKernelGGL((k_expand<GATE_HASH, 0>), dim3(grid), dim3(wg_threads), lds, s, *a);
  } else {
    if (refill) hipLaunchKernelGGL((k_expand<GATE_COUNT, 1>), dim3(grid), dim3(wg_threads), lds, s, *a);
    else hipLaunchKernelGGL((k_expand<GATE_COUNT, 0>), dim3(grid), dim3(wg_threads), lds, s, *a);
  }
  return (int)hipGetLastError();
}

int e13k_launch_batch(const E13BatchArgs *a, int grid, int wg_threads, int gate, hipStream_t s) {
  const size_t lds = e13k_lds_bytes(wg_threads, gate);
  if (gate == GATE_HASH) hipLaunchKernelGGL((k_batch<GATE_HASH>), dim3(grid), dim3(wg_threads), lds, s, *a);
  else hipLaunchKernelGGL((k_batch<GATE_COUNT>), dim3(grid), dim3(wg_threads), lds, s, *a);
  return (int)hipGetLastError();
}

} // extern "C"
