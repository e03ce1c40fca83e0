// e13_kernels.hip -- gfx950 kernels for the TNF enumeration + simulate-and-filter
// hot path (reference: src/e13search.c run()/explore(), src/e13gpu.py expand()).
//
// One machine per lane of a 64-wide wavefront.
//
// LDS, per wave:
//   rows  16 x 64 words at a 4096-byte aligned base: the word of state s for
//         lane l holds the two rule entries (symbol 0 in bits 0..15, symbol 1 in
//         bits 16..31).  An entry is  write | dir << 1 | ones change << 2 |
//         next << 8 (enc_move), or 0x8F00 | kind << 12 for a cell that cannot
//         be stepped (kind 1 halt, 2 unread editable cell,
//         3 free cell not chosen yet).  Because the row base has
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

// LDS words by absolute LDS byte address.  All kernels here use dynamic LDS
// only, which starts at LDS address 0 (checked at kernel start), so offsets
// computed from the start of the dynamic area are absolute addresses and need
// no base added on every access.
typedef __attribute__((address_space(3))) uint32_t lds_word_t;
__device__ __forceinline__ uint32_t lds_ld(const char *sb, int off) {
  (void)sb;
  return *(const lds_word_t *)(size_t)(uint32_t)off;
}
__device__ __forceinline__ void lds_st(char *sb, int off, uint32_t v) {
  (void)sb;
  *(lds_word_t *)(size_t)(uint32_t)off = v;
}
__device__ __forceinline__ void lds_xor(int off, uint32_t v) {
  __hip_atomic_fetch_xor((lds_word_t *)(size_t)(uint32_t)off, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}
__device__ __forceinline__ uint32_t lds_base(const void *p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void *)p;
}

// The 16-bit rule entry of the cell read on symbol `sym`: write | dir << 1 |
// (write - sym) << 2 (three-bit signed, bits 2..4: the ones-count change of the
// step, whose low bit is the tape flip; bits 3 and 4 are both its sign, so a
// two- or a three-bit signed extract at bit 2 reads it) | next << 8; or
// 0x8F00 | kind << 12 for a halt,
// an unread editable cell or a free cell (bit 15: a cell that cannot be
// stepped, so the step loop finds it in the sign of entry << 16).
// The entry of a move code (0..59).
__device__ __forceinline__ uint32_t enc_move(int c, uint32_t sym) {
  const uint32_t d = ((uint32_t)(c & 1) - sym) & 7u; // the change as a 3-bit two's complement field
  return (uint32_t)(c & 3) | (d << 2) | ((uint32_t)(c >> 2) << 8);
}
__device__ __forceinline__ uint32_t enc_entry(int c, uint32_t sym) {
  if (c < 60) return enc_move(c, sym);
  uint32_t kind = c == E13_C_HALT ? 1u : (c == E13_C_EDIT ? 2u : 3u);
  return 0x8F00u | (kind << 12);
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
  int q, row, t, tstop, tcap, dlcap;
  int ones;
  // flags: bit 0 abs_ok, bit 1 rel_ok, bits 2..4 cs[0..2] (the counter hash
  // at that class is set), bit 5 (F_DM) the last step changed `ones` (a cell
  // at q >= 61 flipped), or set when not known.  Before stage 1 the gate has
  // no state or head test, and the block's exactness can change only with
  // `ones`; the step loop's SKIP form therefore tests the stage-1 gate only
  // after a step with F_DM (the gate of a configuration reached without such
  // a step was already settled), so a machine whose two ones sit in the wrong
  // cells does not leave the loop at every step.
  int nt, stage, st0, q0, flags;
  uint64_t hb, tgt;
  uint64_t ch0, ch1, ch2;
};

#define F_DM 32

struct Ctx {
  char *sb;
  int rbl, tbl;       // row and tape byte bases of this lane
  int keyb;           // byte base of the per-workgroup block keys (hash gate)
  const uint64_t *key;
  const uint64_t *tgt;
  const int32_t *dl;
  int dlb;            // >= 0: LDS byte address of the stage deadlines dl[0..9] (used instead of dl)
  int ckb;            // LDS byte address of the counter keys: KEY[q+1] for q 0..60 (CKEY_BYTES)
  int cbl;            // >= 0: LDS byte address of this lane's table codes (8 words, a column of stride 256),
                      // kept beside the rows by the persistent search; -1: decode the rows
  int tw;             // tape words of a lane in LDS (E13_TW: the whole window)
  int qlim;           // largest head position the lane's LDS tape holds (32 * tw - 1; E13_QMAX for the whole
                      // window): a machine past it is paused as a resumable node (EV_CAP) for a launch with
                      // the whole window
  int mvb;            // LDS byte address of word 0, lane 0, of the lone-lane memo's value words (a column of
                      // 16 words per idle lane, stride 256)
};
#define CKEY_BYTES 512 // 64 u64 counter keys per workgroup (61 used)

__device__ __forceinline__ int lane_deadline(const Lane &L, const Ctx &c, int n) {
  if (L.dlcap >= 0) return L.dlcap;
  return c.dlb >= 0 ? (int)lds_ld(c.sb, c.dlb + 4 * n) : c.dl[n];
}

// The gate's targets for stage nt, derived where a round or a slow step needs
// them (so they hold no registers across the persistent loop): the ones count
// 2^nt and the stage-1 state's row address and the heads of either start kind
// (-1000: that kind is ruled out).
struct GatePos {
  int target, row0, qA, qR;
};
__device__ __forceinline__ GatePos gate_pos(int nt, int st0, int q0, int flags, int rbl) {
  GatePos g;
  g.target = 1 << nt;
  g.row0 = rbl + st0 * 256;
  g.qA = (flags & 1) ? q0 : -1000;
  g.qR = (flags & 2) ? q0 - 2 + (1 << nt) : -1000;
  return g;
}

// The step count at which target stage nt's deadline/cap test is due (after a
// load or an accepted stage).
__device__ __forceinline__ void arm_stage(Lane &L, const Ctx &c) {
  L.tstop = min(lane_deadline(L, c, L.nt) + 1, L.tcap);
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
// (q 0..60, KEY index q+1), e13search.c:170.  The 61 keys are in LDS.
typedef __attribute__((address_space(3))) uint64_t lds_u64_t;
__device__ __forceinline__ uint64_t counter_hash(const Ctx &c) {
  uint32_t w0 = lds_ld(c.sb, c.tbl);
  uint32_t w1 = lds_ld(c.sb, c.tbl + 256) & 0x1FFFFFFFu;
  uint64_t h = 0;
  while (w0) {
    int b = __builtin_ctz(w0);
    h ^= *(const lds_u64_t *)(size_t)(uint32_t)(c.ckb + 8 * b);
    w0 &= w0 - 1;
  }
  while (w1) {
    int b = __builtin_ctz(w1);
    h ^= *(const lds_u64_t *)(size_t)(uint32_t)(c.ckb + 8 * (b + 32));
    w1 &= w1 - 1;
  }
  return h;
}

// The counter keys into LDS (the caller synchronises the workgroup after).
__device__ __forceinline__ void fill_ckeys(int ckb, const uint64_t *key) {
  for (int q = threadIdx.x; q < CKEY_BYTES / 8; q += blockDim.x)
    *(lds_u64_t *)(size_t)(uint32_t)(ckb + 8 * q) = q <= 60 ? key[q + 1] : 0ull;
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

// One step in the reference's exact order (e13search.c:136-175): stage gate,
// deadline or launch cap, read, unbound/halt stop, write, move, window, t++.
// The slow path behind the step loop: it always ends the machine or makes the
// step, so a lane never stalls on a configuration.
template <int GATE>
__device__ __forceinline__ int step_once(Lane &L, const Ctx &c, int &evcell) {
  const bool hit = GATE == GATE_COUNT ? L.ones == (1 << L.nt) : L.hb == L.tgt;
  if (hit) {
    const GatePos g = gate_pos(L.nt, L.st0, L.q0, L.flags, c.rbl);
    const bool pre = GATE == GATE_HASH || ((L.nt == 1) | ((L.row == g.row0) & ((L.q == g.qA) | (L.q == g.qR))));
    if (pre) {
      const int ev = stage_event<GATE>(L, c);
      if (ev) return ev;
    }
  }
  if (L.t >= L.tstop) return L.t > lane_deadline(L, c, L.nt) ? EV_DEAD : EV_CAP; // deadline (e13search.c:162) or cap
  const uint32_t qb = (uint32_t)L.q & 31u;
  const int wa = ((L.q >> 5) << 8) + c.tbl;
  uint32_t word = lds_ld(c.sb, wa);
  const uint32_t pair = lds_ld(c.sb, L.row);
  const uint32_t sym = __builtin_amdgcn_ubfe(word, qb, 1u);
  const uint32_t e = pair >> (sym << 4);
  const uint32_t nx = e & 0x0F00u;
  if (nx == 0x0F00u) {
    if (((e >> 12) & 3u) == 1u) return EV_DEAD; // halt read (e13search.c:166)
    evcell = ((L.row - c.rbl) >> 7) + (int)sym;  // first read of an unbound cell (e13search.c:164)
    return EV_STOP;
  }
  const uint32_t w = e & 1u;
  word ^= (w ^ sym) << qb;
  lds_st(c.sb, wa, word);
  if (GATE == GATE_COUNT) {
    const int dm = L.q >= E13_QBLK ? (int)w - (int)sym : 0;
    L.ones += dm;
    L.flags = (L.flags & ~F_DM) | (dm != 0 ? F_DM : 0);
  } else {
    const uint64_t kb = *(const uint64_t *)(c.sb + c.keyb + L.q * 8);
    L.hb ^= kb & (0ull - (uint64_t)(w ^ sym));
  }
  L.q += (int)(e & 2u) - 1;
  if ((unsigned)L.q > (unsigned)E13_QMAX) return EV_DEAD; // off the window; not counted (e13search.c:173)
  L.row = (int)(nx | (uint32_t)c.rbl);
  L.t++;
  return EV_NONE;
}

// Up to `slice` steps.  Returns an event; for EV_STOP the cell is in evcell.
//
// Every step has a single exit test, made on the state before the step
// changes anything: the head left the window on the previous step, the stage
// gate may fire (ones count and stage-1 state/head prefilter, or the block
// hash), the deadline/cap is due, or the cell read is unbound or a halt.  Any
// of them ends the loop and step_once settles the step exactly; everything
// else runs branch-free.  The loop is tested at its bottom: a step's reads and
// exit test are made at the end of the previous iteration, so the loop is one
// block with one exit (no separate masked body per iteration).  A lane that
// ran the whole round has t = its start count + slice; one that left early
// has fewer.
//
// SKIP: the stage-1 gate is tested only after a step that flipped a cell at
// q >= 61 (F_DM in Lane.flags).
// CUT: the round also ends, for every lane still running, once fewer than
// min_live lanes of the wave are still in the loop; such a lane makes its due
// step in step_once (the exact slow path: EV_NONE unless that step ends the
// machine) and goes on next round, while the wave refills its idle lanes.
// The step loop's address and entry arithmetic, written with the operations
// gfx950 issues at twice the rate of three-operand ones (add, and, or, xor and
// right shifts; left shifts, bit-field extracts, compares and selects take the
// full four cycles per wave):
// tape word (q >> 5) of the column tbl: (q & ~31) * 8 + tbl, an and and one
// shift-add (q & ~31 kept opaque, so it is not folded back into a shift and a
// mask of q << 3).  q = -1 gives word -1, an address inside the workgroup's LDS.
__device__ __forceinline__ int tape_word_addr(int q, int tbl) {
  uint32_t qa = (uint32_t)q & ~31u;
  asm("" : "+v"(qa));
  return (int)((qa << 3) + (uint32_t)tbl);
}
// The entry of the symbol at q: (word >> q) << 4 has the symbol at bit 4 and
// zeros below it, so a shift of the pair by its low five bits (the hardware's
// shift masks them) selects bits 0..15 or 16..31.
__device__ __forceinline__ uint32_t entry_of(uint32_t pair, uint32_t word, int q) {
  const uint32_t sh = (word >> ((uint32_t)q & 31u)) << 4;
  uint32_t e;
  asm("v_lshrrev_b32 %0, %1, %2" : "=v"(e) : "v"(sh), "v"(pair));
  return e;
}
__device__ __forceinline__ uint64_t pos_key(int row, int q) { return ((uint64_t)(uint32_t)row << 32) | (uint32_t)q; }
template <int GATE, bool CUT, bool SKIP>
__device__ __forceinline__ int run_slice(Lane &L, const Ctx &c, int slice, int min_live, int &evcell) {
  const int t_end = L.t + slice;
  const int t_lim = min(L.tstop, t_end); // the round's end or the deadline/cap test, whichever comes first
  const bool nt1 = L.nt == 1;
  // steps left before the deadline/cap test or the round's end: the loop runs
  // while left >= 0 (so t = t_lim - 1 - left)
  int left = t_lim - 1 - L.t;
  // the gate's state and head tests as one 64-bit compare each (row address, head)
  const GatePos g = gate_pos(L.nt, L.st0, L.q0, L.flags, c.rbl);
  const uint64_t pA = pos_key(g.row0, g.qA), pR = pos_key(g.row0, g.qR);
  int dmv = L.flags & F_DM; // nonzero: the last step changed the ones count
  // The exit test in one sign bit: (e << 16) is negative for a cell that
  // cannot be stepped (halt, unread), q | (qlim - q) for a head outside the
  // window or past the lane's LDS tape, left for the deadline/cap or round
  // end; the stage gate is tested beside it.
  // q may be one step outside here (q = -1 or qlim + 1): the address is then
  // word -1 or tw, inside the workgroup's LDS, and the word read is not used.
  int wa = tape_word_addr(L.q, c.tbl);
  uint32_t word = lds_ld(c.sb, wa);
  uint32_t pair = lds_ld(c.sb, L.row);
  uint32_t e = entry_of(pair, word, L.q);
  bool stop;
  {
    bool gate;
    if (GATE == GATE_COUNT) {
      const uint64_t pk = pos_key(L.row, L.q);
      gate = (L.ones == g.target) & ((SKIP ? nt1 & (dmv != 0) : nt1) | (pk == pA) | (pk == pR));
    } else {
      gate = L.hb == L.tgt;
    }
    stop = ((int)((e << 16) | (uint32_t)left) < 0) | ((uint32_t)L.q > (uint32_t)c.qlim) | gate;
  }
  while (!stop) {
    // the step
    const uint32_t flip = __builtin_amdgcn_ubfe(e, 2u, 1u); // the write changes the cell
    // flipped in LDS by an atomic xor with no result: the word is read again
    // for the next step, so the register copy needs no update
    lds_xor(wa, flip << ((uint32_t)L.q & 31u));
    if (GATE == GATE_COUNT) {
      // write - read at q >= 61, else 0: a signed extract of width 3 or 0, the
      // width being the top two bits of 60 - q (both set exactly when q > 60)
      const int dm = __builtin_amdgcn_sbfe((int)e, 2u, (uint32_t)(E13_QBLK - 1 - L.q) >> 30);
      L.ones += dm;
      if (SKIP) dmv = dm;
    } else {
      const uint64_t kb = *(const uint64_t *)(c.sb + c.keyb + L.q * 8);
      L.hb ^= kb & (0ull - (uint64_t)flip);
    }
    L.q += (int)(e & 2u) - 1;
    L.row = (int)((e & 0x0F00u) | (uint32_t)c.rbl);
    --left;
    // the next step's reads and exit test
    wa = tape_word_addr(L.q, c.tbl);
    word = lds_ld(c.sb, wa);
    pair = lds_ld(c.sb, L.row);
    e = entry_of(pair, word, L.q);
    bool gate;
    if (GATE == GATE_COUNT) {
      const uint64_t pk = pos_key(L.row, L.q);
      gate = (L.ones == g.target) & ((SKIP ? nt1 & (dmv != 0) : nt1) | (pk == pA) | (pk == pR));
    } else {
      gate = L.hb == L.tgt;
    }
    stop = ((int)((e << 16) | (uint32_t)left) < 0) | ((uint32_t)L.q > (uint32_t)c.qlim) | gate;
    // with too few lanes left in the loop (exec is wave-uniform), every lane
    // leaves it; a lane whose step was due makes it in step_once below
    if (CUT) stop |= __popcll(__builtin_amdgcn_read_exec()) < min_live;
  }
  if (SKIP) L.flags = (L.flags & ~F_DM) | (dmv != 0 ? F_DM : 0);
  const int t = t_lim - 1 - left;
  L.t = t;
  if (t == t_end) return EV_NONE; // a full round (a deadline due exactly now is tested next round)
  if ((unsigned)L.q > (unsigned)E13_QMAX) { // the last step left the window: it does not count
    L.t -= 1;
    return EV_DEAD;
  }
  if (L.q > c.qlim) return EV_CAP; // past the lane's LDS tape (inside the window): paused, goes on wide
  return step_once<GATE>(L, c, evcell);
}

// run_slice for a wave with exactly one running lane, sl: the same steps,
// exit test and hand-off to step_once, run by the whole wave in wave-uniform
// control flow (every lane active) so that every value of the loop is uniform
// and it runs on the scalar unit without exec bookkeeping.  Lane sl's state is
// read once (v_readlane), its 16 rule rows sit in lanes 0..15 of one vector
// register and are read at the state index (v_readlane with a scalar lane
// select), the current tape word is cached (written back when the head
// crosses into another word), and the rare steps are settled by step_once on
// lane sl alone.  A lone lane is bound by the latency of one step, which this
// cuts; ones-count gate only.
__device__ __forceinline__ int solo_uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int lane_of(int v, int sl) { return __builtin_amdgcn_readlane(v, sl); }

// Block macro-steps of the lone lane: the run of the machine inside one
// aligned 16-cell block of tape, up to the step that leaves it or for at most
// SOLO_STAY steps, is a pure function of (state, block contents, entry cell).
// It is memoised in a tagged table of 1024 slots hashed from that key (24
// bits), kept in 32 words of the wave's idle lanes (free while one lane runs
// alone): slot s has its tag (1 << 31 | key) in tape word s >> 6 and its value
// in word s >> 6 of the value area (Ctx.mvb: tape words 16..31, or the rule
// rows when the lane's tape is shorter than 32 words) of lane s & 63 (lane
// sl's slots move to lane (s & 63) ^ 32).  value = new contents | state after << 16 | exit right << 20
// | code << 21, where code 1..SOLO_STAY: the run leaves the block after that
// many steps, on the side the right bit names; code SOLO_STAY+1+p: it is still
// inside after SOLO_STAY steps, at block cell p (a machine caught in a loop
// inside a few cells); code 0: the run reads a halt or an unbound cell.
#define SOLO_MEMO_ROWS 16
#define SOLO_STAY 495

// Run by the whole wave before a lone-lane round: when the lone lane or the
// wave's fill count changed since the last such round (a refill rewrites idle
// lanes' tape words, where the memo lives), the idle lanes clear their tag
// words, so every tag present belongs to the machine now running alone.
__device__ __forceinline__ void solo_memo_reset(bool active, const Ctx &c, uint64_t act, uint64_t fills,
                                                uint64_t &last) {
  const int sl = __ffsll((long long)act) - 1;
  const uint64_t who = (fills << 6) | (uint64_t)sl;
  if (who == last) return;
  last = who;
  if (!active)
    for (int w = 0; w < SOLO_MEMO_ROWS; ++w) lds_st(c.sb, c.tbl + w * 256, 0u);
}

__device__ __forceinline__ int run_solo(Lane &L, const Ctx &c, int nsteps_in, int &evcell, const int sl) {
  const int lane = (int)(threadIdx.x & 63u);
  const int nsteps = solo_uni(nsteps_in);
  const int tbl = lane_of(c.tbl, sl), rbl = lane_of(c.rbl, sl);
  const int tbase = tbl - sl * 4; // the wave's tape area: word w of lane j at tbase + w*256 + j*4
  int q = lane_of(L.q, sl);
  int st = (lane_of(L.row, sl) - rbl) >> 8;
  const int t_round = lane_of(L.t, sl);
  if ((unsigned)q > (unsigned)E13_QMAX) { // a run_slice round ended on the step that left the window
    if (lane == sl) L.t = t_round - 1;    // that step does not count
    return EV_DEAD;
  }
  if (q > c.qlim) return EV_CAP; // a round ended on the step past the lane's LDS tape: paused, goes on wide
  int left = lane_of(L.tstop, sl) - t_round;
  GatePos gs = gate_pos(lane_of(L.nt, sl), lane_of(L.st0, sl), lane_of(L.q0, sl), lane_of(L.flags, sl), 0);
  int target = gs.target, qA = gs.qA, qR = gs.qR;
  int st0r = lane_of(L.st0, sl);
  bool nt1 = lane_of(L.nt, sl) == 1;
  int ones = lane_of(L.ones, sl);
  // rule row k of lane sl in lane k (k < 16; lanes 16..63 hold copies)
  const uint32_t rowv = lds_ld(c.sb, rbl + (lane & 15) * 256);
  // sweep rules: (state, symbol) pairs whose rule rewrites the symbol read and
  // keeps the state; bit `state` of sw0 (symbol 0) or sw1 (symbol 1)
  const uint32_t kv = (uint32_t)(lane & 15) << 8, e0v = rowv & 0xFFFFu, e1v = rowv >> 16;
  const uint32_t sw0 = (uint32_t)__ballot((lane < 16) & ((e0v & 0x0F00u) == kv) & ((e0v & 1u) == 0u));
  const uint32_t sw1 = (uint32_t)__ballot((lane < 16) & ((e1v & 0x0F00u) == kv) & ((e1v & 1u) == 1u));
  int wi = q >> 5;
  uint32_t word = (uint32_t)solo_uni((int)lds_ld(c.sb, (wi << 8) + tbl));
  int n_eff = min(nsteps, left); // the deadline/cap test is due at r == left
  int r = 0, why = 0;            // why = 2: left the window
  for (; r < n_eff; ++r) {
    const uint32_t qb = (uint32_t)q & 31u;
    const uint32_t sym = (word >> qb) & 1u;
    const uint32_t e = (uint32_t)lane_of((int)rowv, st) >> (sym << 4);
    const uint32_t nx = e & 0x0F00u;
    if ((((sym ? sw1 : sw0) >> st) & 1u) != 0u) {
      // a run of equal cells under a rule that rewrites them and keeps the state:
      // each cell of the run is one step that changes nothing but the head, and
      // with the ones count off its target no gate can fire inside it.  Take the
      // run up to the end of this tape word, the window edge and the deadline.
      const uint32_t diff = word ^ (0u - sym); // bits that differ from sym
      int k;
      bool to_end; // the run reaches the end of this tape word
      if (e & 2u) { // right: cells qb, qb+1, ...
        const uint32_t up = diff >> qb;
        to_end = up == 0u;
        k = up ? __builtin_ctz(up) : 32 - (int)qb;
      } else { // left: cells qb, qb-1, ...
        const uint32_t dn = diff << (31u - qb);
        to_end = dn == 0u;
        k = dn ? __builtin_clz(dn) : (int)qb + 1;
      }
      if (to_end) { // it may go on: the following words of the tape, one per lane, in one read
        const int wj = wi + ((e & 2u) ? 1 + lane : -1 - lane);
        const uint32_t fill = 0u - sym;
        uint32_t wv = ~fill; // past the tape: the run ends there
        if (wj >= 0 && wj < c.tw) wv = lds_ld(c.sb, (wj << 8) + tbl);
        const uint64_t mism = __ballot(wv != fill);
        const int j = __builtin_ctzll(mism); // the first word where the run ends (always one)
        const uint32_t wd = (uint32_t)__builtin_amdgcn_readlane((int)wv, j) ^ fill;
        k += 32 * j + ((e & 2u) ? __builtin_ctz(wd) : __builtin_clz(wd));
      }
      k = (e & 2u) ? min(k, c.qlim - q) : min(k, q);
      k = min(k, n_eff - r);
      if (ones == target) { // the gate can fire only in the stage-1 state at qA or qR (or anywhere at nt 1)
        if (nt1 | ((st == st0r) & ((q == qA) | (q == qR)))) k = 0;
        else if (st == st0r) { // stop on the first gate cell ahead; the next iteration tests it
          if (e & 2u) {
            if (qA > q) k = min(k, qA - q);
            if (qR > q) k = min(k, qR - q);
          } else {
            if (qA < q) k = min(k, q - qA);
            if (qR < q) k = min(k, q - qR);
          }
        }
      }
      if (k > 0) {
        q += (e & 2u) ? k : -k;
        r += k - 1; // the loop's ++r counts the last cell
        if ((q >> 5) != wi) {
          lds_st(c.sb, (wi << 8) + tbl, word);
          wi = q >> 5;
          word = (uint32_t)solo_uni((int)lds_ld(c.sb, (wi << 8) + tbl));
        }
        continue;
      }
    }
    // a block macro-step: the 16 cells of tape around the head, away from the
    // window edges (a step off the window is not counted), with the ones count
    // more than 16 from its target (a 16-cell run moves it by at most 16, so
    // no gate can fire) or neither gate cell inside the block (the head visits
    // only its 16 cells: no gate can fire whatever the ones count)
    const int qn = q & ~15;
    const bool far = (ones - target > 16) | (target - ones > 16);
    const bool no_gate_cell = !nt1 & ((qA >> 4) != (qn >> 4)) & ((qR >> 4) != (qn >> 4));
    if ((qn >= 16) & (qn + 16 <= c.qlim) & (far | no_gate_cell)) {
      const uint32_t nb = qb & ~15u, byt = (word >> nb) & 0xFFFFu, pos = qb & 15u;
      const uint32_t key = ((uint32_t)st << 20) | (byt << 4) | pos;
      const uint32_t slot = (key * 0x9E3779B1u) >> 22;
      int col = (int)(slot & 63u);
      if (col == sl) col ^= 32;
      const int ta = tbase + (int)(slot >> 6) * 256 + col * 4, va = c.mvb + (int)(slot >> 6) * 256 + col * 4;
      const uint32_t tag = 0x80000000u | key;
      const uint32_t t_ld = lds_ld(c.sb, ta), v_ld = lds_ld(c.sb, va);
      uint32_t val = (uint32_t)solo_uni((int)v_ld);
      if ((uint32_t)solo_uni((int)t_ld) != tag) { // not yet known: run the block once
        uint32_t v = byt, s2 = (uint32_t)st;
        int p = (int)pos, k2 = 0, right = 0, code = 0;
        while (k2 < SOLO_STAY) {
          const uint32_t b = (v >> p) & 1u;
          const uint32_t e2 = ((uint32_t)lane_of((int)rowv, (int)s2) >> (b << 4)) & 0xFFFFu;
          if ((e2 & 0x0F00u) == 0x0F00u) break; // halt or unbound read: not a plain run
          v = (v & ~(1u << p)) | ((e2 & 1u) << p);
          p += (int)(e2 & 2u) - 1;
          s2 = e2 >> 8;
          ++k2;
          if ((unsigned)p > 15u) {
            right = p > 15;
            code = k2;
            break;
          }
        }
        if (k2 == SOLO_STAY && code == 0) code = SOLO_STAY + 1 + p; // still inside the block
        val = v | (s2 << 16) | ((uint32_t)right << 20) | ((uint32_t)code << 21);
        lds_st(c.sb, ta, tag);
        lds_st(c.sb, va, val);
      }
      const int code = (int)(val >> 21), k3 = min(code, SOLO_STAY);
      if ((code > 0) & (k3 <= n_eff - r)) {
        const uint32_t nv = val & 0xFFFFu;
        word = (word & ~(0xFFFFu << nb)) | (nv << nb);
        // ones are counted for q >= 61: all of a block above, the upper cells of the block at q 48
        const uint32_t cm =
            qn >= E13_QBLK ? 0xFFFFu : (qn + 16 > E13_QBLK ? (0xFFFFu << (E13_QBLK - qn)) & 0xFFFFu : 0u);
        ones += __builtin_popcount(nv & cm) - __builtin_popcount(byt & cm);
        q = code > SOLO_STAY ? qn + (code - SOLO_STAY - 1) : ((val >> 20) & 1u) ? qn + 16 : qn - 1;
        st = (int)((val >> 16) & 15u);
        r += k3 - 1; // the loop's ++r counts the last step
        if ((q >> 5) != wi) {
          lds_st(c.sb, (wi << 8) + tbl, word);
          wi = q >> 5;
          word = (uint32_t)solo_uni((int)lds_ld(c.sb, (wi << 8) + tbl));
        }
        continue;
      }
    }
    if ((nx == 0x0F00u) | (ones == target)) { // rare: settle it exactly in step_once, on lane sl
      if ((nx == 0x0F00u) | nt1 | ((st == st0r) & ((q == qA) | (q == qR)))) {
        lds_st(c.sb, (wi << 8) + tbl, word);
        int ev1 = EV_NONE;
        if (lane == sl) {
          L.q = q;
          L.row = rbl + st * 256;
          L.ones = ones;
          L.t = t_round + r;
          ev1 = step_once<GATE_COUNT>(L, c, evcell);
        }
        ev1 = lane_of(ev1, sl);
        if (ev1 != EV_NONE) return ev1; // the machine ended or stopped; lane sl's state is current
        // a stage start may have been accepted: the gate and deadline registers moved
        q = lane_of(L.q, sl);
        st = (lane_of(L.row, sl) - rbl) >> 8;
        ones = lane_of(L.ones, sl);
        r = lane_of(L.t, sl) - t_round - 1; // the loop's ++r counts the step step_once made
        gs = gate_pos(lane_of(L.nt, sl), lane_of(L.st0, sl), lane_of(L.q0, sl), lane_of(L.flags, sl), 0);
        target = gs.target;
        qA = gs.qA;
        qR = gs.qR;
        st0r = lane_of(L.st0, sl);
        nt1 = lane_of(L.nt, sl) == 1;
        left = lane_of(L.tstop, sl) - t_round;
        n_eff = min(nsteps, left);
        if (q > c.qlim) { // the step went past the lane's LDS tape: paused
          // (step_once wrote the cell in word wi: the cached word is stale)
          word = (uint32_t)solo_uni((int)lds_ld(c.sb, (wi << 8) + tbl));
          r += 1;
          why = 3;
          break;
        }
        wi = q >> 5;
        word = (uint32_t)solo_uni((int)lds_ld(c.sb, (wi << 8) + tbl));
        continue;
      }
    }
    const uint32_t w = e & 1u;
    word ^= (w ^ sym) << qb;
    ones += q >= E13_QBLK ? (int)w - (int)sym : 0;
    q += (int)(e & 2u) - 1;
    if ((unsigned)q > (unsigned)E13_QMAX) { // off the window: dead, and this step does not count
      why = 2;
      break;
    }
    st = (int)(nx >> 8);
    if (q > c.qlim) { // past the lane's LDS tape (the step counts): paused, goes on wide
      r += 1;
      why = 3;
      break;
    }
    if ((q >> 5) != wi) { // into another tape word
      lds_st(c.sb, (wi << 8) + tbl, word);
      wi = q >> 5;
      word = (uint32_t)solo_uni((int)lds_ld(c.sb, (wi << 8) + tbl));
    }
  }
  lds_st(c.sb, (wi << 8) + tbl, word);
  int ev = EV_NONE;
  if (lane == sl) {
    L.q = q;
    L.row = rbl + st * 256;
    L.ones = ones;
    L.t = t_round + r;
    L.flags |= F_DM; // the loop's steps are not tracked: the gate is tested again
    if (why == 2) ev = EV_DEAD;
    else if (why == 3) ev = EV_CAP;
    else if (r >= left) ev = step_once<GATE_COUNT>(L, c, evcell); // the deadline or cap is due
  }
  return lane_of(ev, sl);
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ long long uni64(long long v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((unsigned long long)v >> 32));
  return (long long)(((unsigned long long)hi << 32) | lo);
}

// Core-clock time stamp for the per-phase cycle counters.
__device__ __forceinline__ uint64_t stamp() { return __builtin_readcyclecounter(); }
// The same when the per-phase cycle counts are asked for, else 0 (a clock read
// costs its latency on every phase boundary, about 2% of the search).
__device__ __forceinline__ uint64_t stamp_if(int on) { return on ? __builtin_readcyclecounter() : 0ull; }

// Wave event log of the persistent search (when a.trace is set): lane 0 appends
// (constant-rate real time << 16 | aux << 8 | event) to the wave's own region.
enum { TR_SEEDED = 1, TR_TAKE = 2, TR_RUNNING = 3, TR_IDLE = 4, TR_EXIT = 5, TR_GIVE = 6, TR_ROUND = 7 };
#define TRACE_EV2(ev, aux)                                                                            \
  do {                                                                                                \
    if (CLOCKS && a.trace != nullptr && lane == 0) {                                                  \
      if (tr_n + 1u < (uint32_t)a.trace_cap)                                                          \
        a.trace[(size_t)wid * (size_t)a.trace_cap + 1u + tr_n] =                                      \
            (__builtin_readsteadycounter() << 16) | ((uint64_t)((aux) & 0xFF) << 8) | (ev);           \
      ++tr_n;                                                                                         \
    }                                                                                                 \
  } while (0)
#define TRACE_EV(ev) TRACE_EV2(ev, 0)

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
    lds_st(c.sb, c.rbl + r * 256, enc_entry(c0, 0u) | (enc_entry(c1, 1u) << 16));
  }
}

// Codes of cells 4i..4i+3 packed little-endian (rows past nq read as free).
__device__ __forceinline__ uint32_t codes_word(const Ctx &c, int nq, int i) {
  if (c.cbl >= 0) return lds_ld(c.sb, c.cbl + i * 256);
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
  c.dlb = -1;
  c.cbl = -1;
  c.tw = E13_TW;
  c.qlim = E13_QMAX;
  c.mvb = c.tbl - lane * 4 + SOLO_MEMO_ROWS * 256; // memo values in tape words 16..31
  if (lds_base(smem) != 0u) { // the absolute LDS addressing needs the dynamic area at address 0
    if (threadIdx.x == 0) atomicOr(&a.ctr->overflow, 4u);
    return;
  }
  // per-wave stage histogram: 16 words after the tapes and block keys
  uint32_t *whist = (uint32_t *)(c.sb + c.keyb + (GATE == GATE_HASH ? KEYB_WORDS * 8 : 0)) + wave * 16;
  if (lane < 16) whist[lane] = 0;
  if (GATE == GATE_HASH) fill_keyb(c.sb, c.keyb, a.key);
  c.ckb = c.keyb + (GATE == GATE_HASH ? KEYB_WORDS * 8 : 0) + nwaves * 64;
  fill_ckeys(c.ckb, a.key);
  __syncthreads();

  uint32_t nin = a.nin;
  const E13Node *in = a.in;
  if (a.chain != nullptr) { // launch chain: this iteration's batch is on the device
    nin = a.chain->nin;
    in = a.stack + a.chain->base;
  }
  const uint32_t nitems = nin * a.nslot;
  Lane L;
  L.dlcap = -1;
  int active = 0, t0 = 0, budget = 0, nintro = 0, front = 0, depth = 0, nbound = 0;
  uint32_t seed = 0; // frontier node the lane's machine descends from
  uint64_t my_steps = 0;
  uint32_t my_leaves = 0, my_surv = 0;
  uint64_t w_slots = 0, w_rounds = 0;
  bool exhausted = false;
  bool first = true;
  uint32_t ck = 0, ce = 0; // this wave's claimed item range
  uint64_t cyc0 = 0, cyc1 = 0, cyc2 = 0, ts = stamp();

  for (;;) {
    // ---- refill ----
    if (REFILL || first) {
      uint64_t need = ballot(!active);
      for (int tries = 0; need != 0 && tries < a.refill_tries; ++tries) {
        uint32_t k;
        if (REFILL) {
          // items come from a wave-private range claimed `chunk` at a time:
          // one atomic on the shared cursor per chunk, not per round
          if (ck >= ce) {
            if (exhausted) break;
            uint32_t b0 = 0;
            if (lane == 0) b0 = atomicAdd(&a.ctr->item, a.chunk);
            b0 = (uint32_t)uni((int)b0);
            if (b0 >= nitems || b0 + a.chunk >= nitems) exhausted = true;
            ck = b0;
            ce = min(b0 + a.chunk, nitems);
            if (ck >= ce) break;
          }
          const int cnt = __popcll(need);
          const uint32_t give = min((uint32_t)cnt, ce - ck);
          const uint32_t rk = rank_in(need);
          k = rk < give ? ck + rk : nitems;
          ck += give;
        } else {
          if (tries > 0) break; // one item per lane: never take the same item twice
          k = blockIdx.x * blockDim.x + threadIdx.x;
          exhausted = true;
        }
        if (!active && k < nitems) {
          const uint32_t ni = k / a.nslot, slot = k - ni * a.nslot;
          const E13Node *nd = in + ni;
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
            L.hb = GATE == GATE_HASH ? nd->hb : 0ull;
            L.ch0 = nd->ch[0];
            L.ch1 = nd->ch[1];
            L.ch2 = nd->ch[2];
            L.tcap = L.t + a.cap;
            L.flags |= F_DM; // not known: the first step tests the gate
            arm_stage(L, c);
            t0 = L.t;
            budget = nb;
            nintro = nin2;
            front = nd->prob;
            seed = nd->seed;
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
      if (exhausted && ck >= ce) break;
      continue;
    }
    w_rounds++;
    w_slots += (uint64_t)__popcll(act);
    uint64_t tn = stamp();
    cyc0 += tn - ts;
    ts = tn;
    // ---- a round of steps ----
    int ev = EV_NONE, evcell = -1;
    if (active) ev = run_slice<GATE, false, false>(L, c, a.slice, 0, evcell);
    tn = stamp();
    cyc1 += tn - ts;
    ts = tn;
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
          o->prob = front;
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
          o->prob = front;
          o->depth = depth;
          o->nbound = nbound;
          o->next_slot = 0;
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
            v4.w = (4 * w4 + 3 < E13_TW) ? lds_ld(c.sb, c.tbl + (4 * w4 + 3) * 256) : seed;
            dst[w4] = v4;
          }
        } else {
          atomicOr(&a.ctr->overflow, 1u);
        }
      }
    }
    tn = stamp();
    cyc2 += tn - ts;
    ts = tn;
    if (!REFILL && ballot(active) == 0) break;
  }
  // ---- flush counters ----
  if (lane == 0) {
    atomicAdd((unsigned long long *)&a.ctr->cyc[0], (unsigned long long)cyc0);
    atomicAdd((unsigned long long *)&a.ctr->cyc[1], (unsigned long long)cyc1);
    atomicAdd((unsigned long long *)&a.ctr->cyc[2], (unsigned long long)cyc2);
  }
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
template <int GATE, int REFILL>
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
  c.dlb = -1;
  c.cbl = -1;
  c.tw = E13_TW;
  c.qlim = E13_QMAX;
  c.mvb = c.tbl - lane * 4 + SOLO_MEMO_ROWS * 256; // memo values in tape words 16..31
  if (lds_base(smem) != 0u) { // the absolute LDS addressing needs the dynamic area at address 0
    if (threadIdx.x == 0) atomicOr(&a.ctr->overflow, 4u);
    return;
  }
  if (GATE == GATE_HASH) fill_keyb(c.sb, c.keyb, a.key);
  c.ckb = c.keyb + (GATE == GATE_HASH ? KEYB_WORDS * 8 : 0) + nwaves * 64;
  fill_ckeys(c.ckb, a.key);
  __syncthreads();
  Lane L;
  int active = 0;
  uint32_t k = 0;
  uint64_t my_steps = 0, w_slots = 0, w_rounds = 0;
  uint64_t solo_last = ~0ull, mach = 0; // mach: lane slots this wave has handed out (wave-uniform)
  bool exhausted = false;
  uint32_t ck = 0, ce = 0; // this wave's claimed table range
  if (!REFILL) { // plain form: one table per lane, no refill
    ck = blockIdx.x * blockDim.x + (threadIdx.x & ~63u);
    ce = min(ck + 64u, (uint32_t)a.n);
    exhausted = true;
  }
  for (;;) {
    uint64_t need = ballot(!active);
    for (int tries = 0; need != 0 && tries < 2; ++tries) {
      if (ck >= ce) {
        if (exhausted) break;
        // tables come from a wave-private range claimed `chunk` at a time (one
        // atomic on the shared cursor per chunk); chunk 0: as many as lanes are idle
        const uint32_t want = a.chunk ? a.chunk : (uint32_t)__popcll(need);
        uint32_t b0 = 0;
        if (lane == 0) b0 = atomicAdd(&a.ctr->item, want);
        b0 = (uint32_t)uni((int)b0);
        if (b0 >= (uint32_t)a.n || b0 + want >= (uint32_t)a.n) exhausted = true;
        ck = b0;
        ce = min(b0 + want, (uint32_t)a.n);
        if (ck >= ce) break;
      }
      const uint32_t give = min((uint32_t)__popcll(need), ce - ck);
      const uint32_t rk = rank_in(need);
      if (!active && rk < give) {
        k = ck + rk;
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
        L.flags |= F_DM; // not known: the first step tests the gate
        arm_stage(L, c);
        active = 1;
      }
      ck += give;
      mach += give;
      need = ballot(!active);
    }
    const uint64_t act = ballot(active);
    if (act == 0) {
      if (exhausted && ck >= ce) break;
      continue;
    }
    w_rounds++;
    w_slots += (uint64_t)__popcll(act);
    int ev = EV_NONE, evcell = -1;
    // a lone running lane takes the scalar loop (longer rounds: nothing to refill meanwhile
    // once the batch is exhausted)
    const bool solo = GATE == GATE_COUNT && a.solo > 0 && __popcll(act) == 1;
    if (solo) { // the whole wave runs the lone lane's loop
      solo_memo_reset(active, c, act, mach, solo_last);
      ev = run_solo(L, c, exhausted ? a.solo : a.slice, evcell, __ffsll((long long)act) - 1);
      if (!active) ev = EV_NONE;
    } else if (active) {
      ev = run_slice<GATE, false, false>(L, c, a.slice, 0, evcell);
    }
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

// Number the frontier nodes of a persistent search (cost logging): seed i = i.
__global__ void __launch_bounds__(256) k_seed_ids(E13Node *s, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) s[i].seed = i;
}

// ---------------------------------------------------------------------------
// Launch chains of the host-loop forms.  One iteration is k_chain_pop ->
// k_expand (chain mode) -> k_chain_push; the stack cursor stays on the device,
// so a hipGraph of many iterations runs with no host round trip.  An
// iteration that finds the stack empty does nothing.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_chain_pop(E13Chain *ch, E13Counters *ctr) {
  if (threadIdx.x == 0) {
    const uint64_t top = ch->error ? 0ull : ch->top;
    const uint32_t nin = top < (uint64_t)ch->batch ? (uint32_t)top : ch->batch;
    ch->nin = nin;
    ch->base = top - nin;
    if (nin) ch->iters += 1u;
    ctr->item = 0u; // the expansion's work cursor and child count restart (no memset node in a graph)
    ctr->nout = 0u;
  }
}

__global__ void __launch_bounds__(256) k_chain_push(E13Chain *ch, E13Node *stack, uint64_t stack_cap,
                                                    const E13Node *out, uint32_t out_cap, const E13Counters *ctr) {
  const uint32_t nout = ctr->nout, nin = ch->nin;
  const uint64_t base = ch->base;
  if (nin == 0u) return;
  if (ctr->overflow != 0u || nout > out_cap || base + nout > stack_cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) ch->error = (ctr->overflow & 4u) ? 2u : 1u;
    return;
  }
  const uint4 *src = (const uint4 *)out;
  uint4 *dst = (uint4 *)(stack + base);
  const uint64_t n16 = (uint64_t)nout * (sizeof(E13Node) / 16);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) { // read only by the next kernel in stream order
    ch->top = base + nout;
    if (base + nout > ch->peak) ch->peak = base + nout;
  }
}

// ---------------------------------------------------------------------------
// Persistent whole search (one launch per search).  Each wave owns a private
// depth-first node stack in HBM.  The node on top is staged once in LDS and
// its child slots are handed to idle lanes; a lane that stops at an unbound
// cell pushes its machine as a new node.  Idle waves take nodes from a
// global pool that busy waves feed with their oldest nodes.  The search ends
// when no wave holds work and the pool is empty; every wait is bounded and
// any failure sets an abort flag that all waves watch.
// ---------------------------------------------------------------------------
namespace {

#define START_LIMIT 2000000ull // the residency check's wait: 20 ms of the 100 MHz constant clock
#define STG_WORDS 120 // staged node (76) + rows as stored (16) + rows with edits exhausted (16)
                      // + codes with edits exhausted (8) + 4 spare
#define STG_W_CODESB 108
#define CODES_BYTES (8 * 256) // per wave: each lane's 32 table codes, word i at i * 256 + lane * 4

// Per-workgroup copy of the problems in LDS (read by the refill, which would
// otherwise wait on global loads): problem k at word k * PT_WORDS holds its 32
// base codes as bytes (words 0..7), then nq, nslot, sym_lo, sym_hi; the stage
// deadlines dl[0..9] (the same for every problem) follow the last problem.
#define PT_WORDS 16
#define PT_DL_WORDS 16
__device__ __forceinline__ int pt_word(int ptb, int k, int w) {
  return (int)*(const lds_word_t *)(size_t)(uint32_t)(ptb + (k * PT_WORDS + w) * 4);
}
__device__ __forceinline__ int pt_base(int ptb, int k, int i) {
  return (int)*(const __attribute__((address_space(3))) int8_t *)(size_t)(uint32_t)(ptb + k * PT_WORDS * 4 + i);
}

// This lane's LDS tape column and tape length as opaque values (when the
// length is not a constant): a test `w < tw` per word on the plain values is
// hoisted out of the persistent loop, and the 43 test results held across it
// spill; an opaque column base lets the words' offsets fold into the reads.
struct LaneTape {
  int tb, tw;
};
__device__ __forceinline__ LaneTape lane_tape(const Ctx &c) {
  LaneTape t{c.tbl, c.tw};
  asm volatile("" : "+v"(t.tb));
  if (!__builtin_constant_p(t.tw)) asm volatile("" : "+s"(t.tw));
  __builtin_assume(t.tw >= E13_TW_MIN && t.tw <= E13_TW);
  return t;
}

// Words W.. tw-1 of a loaded lane tape (tw at most E13_TW), word w at tb + w * 256.
template <int W> __device__ __forceinline__ void tape_tail(const Ctx &c, int tb, const uint32_t (&tv)[44], int tw) {
  if constexpr (W < E13_TW) {
    if (tw > W) {
      lds_st(c.sb, tb + W * 256, tv[W]);
      tape_tail<W + 1>(c, tb, tv, tw);
    }
  }
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}
__device__ __forceinline__ void wave_mem_sync() { // this wave's global stores visible to its own later loads
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

// A node's child slots still to hand out, as stored in E13Node.next_slot:
// first | end << 16 (end 0 = all of the problem's slots).
__device__ __forceinline__ int slot_range(int first, int end) { return first | (end << 16); }

__device__ __forceinline__ E13Node *stk_node(const E13SearchArgs &a, int wid, long long idx) {
  // stk_cap is a power of two: the ring slot is the index's low bits
  const uint32_t slot = (uint32_t)wid * (uint32_t)a.stk_cap + ((uint32_t)idx & (uint32_t)(a.stk_cap - 1));
  return a.stacks + slot;
}

__device__ __forceinline__ uint64_t ld_agent64(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent64(uint64_t *p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_agent_i(const int32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Wait until every memory operation this wave issued has completed.  Data
// shared across CUs is written and read only with agent-scope (sc1) accesses,
// which are coherent without cache maintenance, so a completed store is
// visible to any later agent-scope load: a drain before a flag store makes
// the flag a publication of everything before it (no buffer_wbl2/buffer_inv).
__device__ __forceinline__ void vm_drain() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// The search is over (or aborted): every wave's mailbox says so.
// (the fields, not the argument block: a reference to the kernel's argument
// block in a call that is not inlined would make the compiler copy the block
// to scratch memory)
__device__ void broadcast_done(E13Mail *mail, int total_waves, int lane) {
  for (int w = lane; w < total_waves; w += 64) st_agent64(&mail[w].flag, E13_MAIL_DONE);
  vm_drain();
}

// Encode both row sets of the staged node: A as stored, B with unread
// editable cells bound to their base values (a child that spends the last edit).
__device__ void stage_rows(uint32_t *stg, int ptb, int lane) {
  wave_lds_sync();
  const int prob = uni((int)stg[12]);
  const int nq = pt_word(ptb, prob, 8);
  if (lane < 16) {
    uint32_t A = 0, B = 0;
    if (lane < nq) {
      const uint8_t *tab = (const uint8_t *)(stg + E13_NODE_W_TAB);
      int c0 = tab[2 * lane], c1 = tab[2 * lane + 1];
      A = enc_entry(c0, 0u) | (enc_entry(c1, 1u) << 16);
      const int b0 = pt_base(ptb, prob, 2 * lane), b1 = pt_base(ptb, prob, 2 * lane + 1);
      if (c0 == E13_C_EDIT) c0 = b0 < 0 ? E13_C_HALT : b0;
      if (c1 == E13_C_EDIT) c1 = b1 < 0 ? E13_C_HALT : b1;
      B = enc_entry(c0, 0u) | (enc_entry(c1, 1u) << 16);
    }
    stg[76 + lane] = A;
    stg[92 + lane] = B;
  }
  if (lane < E13_MAXCELLS) { // codes with edits exhausted: unread editable cells keep their base value
    const uint8_t *tab = (const uint8_t *)(stg + E13_NODE_W_TAB);
    int cd = tab[lane];
    if (cd == E13_C_EDIT) {
      const int b = pt_base(ptb, prob, lane);
      cd = b < 0 ? E13_C_HALT : b;
    }
    ((uint8_t *)(stg + STG_W_CODESB))[lane] = (uint8_t)cd;
  }
  wave_lds_sync();
}

// Stage node `nd` (76 words) from global memory into LDS words stg[0..75].
__device__ void stage_node(uint32_t *stg, const uint32_t *nd, int ptb, int lane) {
  const uint32_t v0 = nd[lane];
  const uint32_t v1 = lane < E13_NODE_WORDS - 64 ? nd[64 + lane] : 0u;
  stg[lane] = v0;
  if (lane < E13_NODE_WORDS - 64) stg[64 + lane] = v1;
  stage_rows(stg, ptb, lane);
}

// The newest child as the staged node, without a reload from HBM: words
// 0..23 (the machine's registers) from its own lane, the same words
// write_node stores ...
__device__ __forceinline__ void stage_head(const Ctx &c, uint32_t *stg, const Lane &L, int budget, int nintro, int pend,
                                           int prob, int depth, int nbound) {
  stg[0] = (uint32_t)L.q;
  stg[1] = (uint32_t)((L.row - c.rbl) >> 8);
  stg[2] = (uint32_t)L.t;
  stg[3] = (uint32_t)L.nt;
  stg[4] = (uint32_t)L.stage;
  stg[5] = (uint32_t)L.st0;
  stg[6] = (uint32_t)L.q0;
  stg[7] = (uint32_t)L.flags;
  stg[8] = (uint32_t)L.ones;
  stg[9] = (uint32_t)budget;
  stg[10] = (uint32_t)nintro;
  stg[11] = (uint32_t)pend;
  stg[12] = (uint32_t)prob;
  stg[13] = (uint32_t)depth;
  stg[14] = (uint32_t)nbound;
  stg[15] = 0u;
  stg[16] = (uint32_t)L.hb;
  stg[17] = (uint32_t)(L.hb >> 32);
  stg[18] = (uint32_t)L.ch0;
  stg[19] = (uint32_t)(L.ch0 >> 32);
  stg[20] = (uint32_t)L.ch1;
  stg[21] = (uint32_t)(L.ch1 >> 32);
  stg[22] = (uint32_t)L.ch2;
  stg[23] = (uint32_t)(L.ch2 >> 32);
}

// ... and words 24..75 (table codes, tape, seed) from lane src's LDS columns
// by the whole wave, one word per lane (lanes 0..51): two reads and a store
// per lane instead of 52 reads and stores by one lane.  Without a codes
// column (c.cbl < 0), lanes 0..7 decode codes word k from rows 2k, 2k+1.
__device__ __forceinline__ void stage_body(const Ctx &c, uint32_t *stg, int src, uint32_t seed, int nq, int lane) {
  const uint32_t sd = (uint32_t)__builtin_amdgcn_readlane((int)seed, src);
  const int nqs = __builtin_amdgcn_readlane(nq, src);
  const LaneTape lt = lane_tape(c);
  const int d = (src - lane) * 4; // lane src's column from this lane's
  const int k = lane - E13_MAXCELLS / 4; // tape word of this lane, if it has one
  uint32_t v = 0u;
  if (k < 0) {
    if (c.cbl >= 0) {
      v = lds_ld(c.sb, c.cbl + d + lane * 256);
    } else { // as codes_word, on lane src's rows (rows past its nq read as free)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int r = 2 * lane + h;
        uint32_t lo = E13_C_FREE, hi = E13_C_FREE;
        if (r < nqs) {
          const uint32_t pr = lds_ld(c.sb, c.rbl + d + r * 256);
          lo = (uint32_t)dec_entry(pr);
          hi = (uint32_t)dec_entry(pr >> 16);
        }
        v |= (lo | (hi << 8)) << (16 * h);
      }
    }
  } else if (k < lt.tw) v = lds_ld(c.sb, lt.tb + d + k * 256);
  else if (k == E13_TW) v = sd;
  if (k < E13_TW_PAD) stg[E13_NODE_W_TAB + lane] = v;
}

// Copy this lane's machine into node record `o` (plain stores).
__device__ __forceinline__ void write_node(const Ctx &c, E13Node *o, const Lane &L, int budget, int nintro, int pend,
                                           int prob, int depth, int nbound, int nq, uint32_t seed) {
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
  o->pend = pend;
  o->prob = prob;
  o->depth = depth;
  o->nbound = nbound;
  o->next_slot = 0;
  o->hb = L.hb;
  o->ch[0] = L.ch0;
  o->ch[1] = L.ch1;
  o->ch[2] = L.ch2;
  uint32_t *tw = (uint32_t *)o->tab;
#pragma unroll
  for (int i = 0; i < E13_MAXCELLS / 4; ++i) tw[i] = codes_word(c, nq, i);
  uint4 *dst = (uint4 *)o->tape;
  const LaneTape lt = lane_tape(c);
#pragma unroll
  for (int w4 = 0; w4 < E13_TW_PAD / 4; ++w4) {
    uint4 v4;
    // words past the lane's LDS tape are blank (the machine never went there); word 43 is the seed
    v4.x = 4 * w4 + 0 < lt.tw ? lds_ld(c.sb, lt.tb + (4 * w4 + 0) * 256) : 0u;
    v4.y = 4 * w4 + 1 < lt.tw ? lds_ld(c.sb, lt.tb + (4 * w4 + 1) * 256) : 0u;
    v4.z = 4 * w4 + 2 < lt.tw ? lds_ld(c.sb, lt.tb + (4 * w4 + 2) * 256) : 0u;
    v4.w = 4 * w4 + 3 < lt.tw ? lds_ld(c.sb, lt.tb + (4 * w4 + 3) * 256) : (4 * w4 + 3 < E13_TW ? 0u : seed);
    dst[w4] = v4;
  }
}

} // namespace

// NARROW: each lane's tape in LDS is a.tw < E13_TW words (cells up to
// q = 32 * a.tw - 1), so more waves fit in a CU.  A machine whose head passes
// that cell is paused as a resumable node (as at a launch cap), and a node
// whose head or written cells lie past it is not loaded: it goes, with its
// slot cursor, to the list a.wide, which the host runs afterwards in a launch
// with the whole window.  When that list is full the launch stops cleanly
// (every wave pauses its machines and leaves), so nothing is lost.  The
// NARROW form is built for four waves per SIMD (128 VGPRs: the step loop
// keeps its registers, the refill and finish spill) and keeps no table codes
// in LDS (they are decoded from the rule rows where a node or a survivor is
// written), so four 4-wave workgroups fit a CU with an 18-word tape.
// k_search's per-lane machine bookkeeping in one word: problem index (6
// bits, < E13_MAXPROB), rows nq (6), symmetric states in use (6), binding
// depth (7) and cells bound (7), each at most E13_MAXCELLS.
__device__ __forceinline__ uint32_t meta_pack(int prob, int nq, int nintro, int depth, int nbound) {
  return (uint32_t)prob | ((uint32_t)nq << 6) | ((uint32_t)nintro << 12) | ((uint32_t)depth << 18) |
         ((uint32_t)nbound << 25);
}
__device__ __forceinline__ int meta_prob(uint32_t m) { return (int)(m & 63u); }
__device__ __forceinline__ int meta_nq(uint32_t m) { return (int)((m >> 6) & 63u); }
__device__ __forceinline__ int meta_nintro(uint32_t m) { return (int)((m >> 12) & 63u); }
__device__ __forceinline__ int meta_depth(uint32_t m) { return (int)((m >> 18) & 127u); }
__device__ __forceinline__ int meta_nbound(uint32_t m) { return (int)(m >> 25); }

// CLOCKS: the per-phase wave clock reads (tuning phase_cycles) and the wave
// event trace (trace_events) are compiled in; the short-tape count-gate search
// is also built without them (k_search), so their accumulators and the trace
// cursor hold no registers in the timed kernel.
template <int GATE, bool NARROW, bool CLOCKS>
__device__ __forceinline__ void search_body(const E13SearchArgs &a) {
  extern __shared__ __align__(16) uint32_t smem[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  const int wid = blockIdx.x * nwaves + wave;
  const int tw = NARROW ? a.tw : E13_TW;
  Ctx c;
  c.sb = (char *)smem;
  c.rbl = wave * ROWS_BYTES + lane * 4;
  c.tbl = nwaves * ROWS_BYTES + wave * tw * 256 + lane * 4;
  c.keyb = nwaves * (ROWS_BYTES + tw * 256);
  c.tw = tw;
  c.qlim = NARROW ? 32 * tw - 1 : E13_QMAX;
  // the lone-lane memo's values: tape words 16..31 of idle lanes, or their rule rows on a shorter tape
  c.mvb = tw >= 2 * SOLO_MEMO_ROWS ? c.tbl - lane * 4 + SOLO_MEMO_ROWS * 256 : wave * ROWS_BYTES;
  c.key = a.key;
  c.tgt = a.tgt;
  c.dl = a.pbs[0].dl;
  if (lds_base(smem) != 0u) { // the absolute LDS addressing needs the dynamic area at address 0
    if (threadIdx.x == 0) atomicMax(&a.sch->error, 4);
    return;
  }
  if (a.nseeds == 0) return; // no work: no wave would ever count busy down to the end mark
  // Every workgroup must be resident at once (seeds are dealt to every wave,
  // and a wave that started only after others left would hold its share and
  // its place in the busy count).  Each workgroup counts itself in and waits,
  // bounded, for all of them; if they do not all arrive in time, or a
  // workgroup arrives after others gave up, every workgroup leaves before any
  // work (error 5) and the host redoes the launch in a shape that fits.
  if (threadIdx.x == 0) {
    int ok = 1;
    atomicAdd(&a.sch->started, 1);
    const uint64_t t_start = __builtin_readsteadycounter();
    for (;;) {
      if (ld_agent_i(&a.sch->error) != 0) { // another workgroup gave up (or failed): leave too
        ok = 0;
        break;
      }
      if (ld_agent_i(&a.sch->started) >= (int)gridDim.x) break;
      if (__builtin_readsteadycounter() - t_start > START_LIMIT) {
        atomicCAS(&a.sch->error, 0, 5);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    smem[0] = (uint32_t)ok;
  }
  __syncthreads();
  const bool resident = smem[0] != 0u;
  __syncthreads(); // smem[0] is wave 0's first rule word: read by all before it is written
  if (!resident) return;
  const int after_keys = c.keyb + (GATE == GATE_HASH ? KEYB_WORDS * 8 : 0);
  uint32_t *stg = (uint32_t *)(c.sb + after_keys) + wave * STG_WORDS;
  unsigned long long *pc = (unsigned long long *)(c.sb + after_keys + nwaves * STG_WORDS * 4); // np x 16 per WG
  for (int i = threadIdx.x; i < a.np * E13_PCNT; i += blockDim.x) pc[i] = 0ull;
  // the problems and the stage deadlines, in LDS
  const int ptb = after_keys + nwaves * STG_WORDS * 4 + a.np * E13_PCNT * 8;
  for (int i = threadIdx.x; i < a.np * PT_WORDS; i += blockDim.x) {
    const int k = i / PT_WORDS, w = i - k * PT_WORDS;
    const E13Problem *pb = a.pbs + k;
    uint32_t v = 0;
    if (w < 8) v = ((const uint32_t *)pb->base)[w];
    else if (w == 8) v = (uint32_t)pb->nq;
    else if (w == 9) v = (uint32_t)pb->nslot;
    else if (w == 10) v = (uint32_t)pb->sym_lo;
    else if (w == 11) v = (uint32_t)pb->sym_hi;
    lds_st(c.sb, ptb + i * 4, v);
  }
  c.dlb = ptb + a.np * PT_WORDS * 4;
  for (int i = threadIdx.x; i < PT_DL_WORDS; i += blockDim.x)
    lds_st(c.sb, c.dlb + i * 4, i <= E13_NS ? (uint32_t)a.pbs[0].dl[i] : 0u);
  c.ckb = c.dlb + PT_DL_WORDS * 4; // 8-byte aligned: every area before it is a multiple of 8 bytes
  fill_ckeys(c.ckb, a.key);
  c.cbl = NARROW ? -1 : c.ckb + CKEY_BYTES + wave * CODES_BYTES + lane * 4;
  if (GATE == GATE_HASH) fill_keyb(c.sb, c.keyb, a.key);
  __syncthreads();

  Lane L;
  L.dlcap = -1;
  L.tcap = 0x7FFFFFFF;
  int active = 0, t0 = 0, budget = 0;
  // the machine's problem index, rows, symmetric states in use, binding depth
  // and cells bound, packed in one register (read where a machine finishes)
  uint32_t meta = 0;
  uint32_t seed = 0; // frontier node the lane's machine descends from
  uint64_t solo_last = ~0ull, mach = 0; // mach: lane slots this wave has handed out (wave-uniform)
  long long top = 0, bot = 0, staged = -1;
  bool stg_checked = false; // NARROW: the staged node was tested against the lane tape
  int cur_slot = 0, cur_end = 0; // the staged node's child slots still to hand out: [cur_slot, cur_end)
  int busy = 0, requested = 0, mail_in = 1, idle_polls = 0;
  uint64_t last_tag = 0; // tag of the last delivery taken from this wave's mailbox
  uint64_t idle_since = 0; // constant-clock time this wave went idle (0: not idle)
  // the frontier from the breadth-first expansion: seed i goes to wave i mod total_waves
  // (the host counted every wave that receives one as busy)
  {
    const int W = a.total_waves;
    const int mine = (int)(a.nseeds / (uint32_t)W) + (wid < (int)(a.nseeds % (uint32_t)W) ? 1 : 0);
    for (int j = 0; j < mine; ++j) {
      const uint32_t *src = (const uint32_t *)(a.seeds + (size_t)wid + (size_t)j * (size_t)W);
      uint32_t *dst = (uint32_t *)stk_node(a, wid, j);
      const uint32_t v0 = src[lane];
      const uint32_t v1 = lane < E13_NODE_WORDS - 64 ? src[64 + lane] : 0u;
      dst[lane] = v0;
      if (lane < E13_NODE_WORDS - 64) dst[64 + lane] = v1;
    }
    top = mine;
    busy = mine > 0 ? 1 : 0;
  }
  uint32_t tr_n = 0;
  if (busy) TRACE_EV(TR_SEEDED);
  int refresh_in = (wid % 8) + 1;
  uint64_t w_slots = 0, w_rounds = 0, w_takes = 0, w_gives = 0;
  uint32_t w_steps = 0;       // this lane's steps since the last fold (below 2^31 + 2^31)
  uint64_t w_steps_wave = 0;  // the wave's folded steps (uniform)
  uint64_t cyc0 = 0, cyc1 = 0, cyc2 = 0, cyc3 = 0, ts = (CLOCKS ? stamp_if(a.phase_cycles) : 0ull), tn;
  bool halt_self = false; // NARROW: the list of wide nodes is full; this wave stops (and so do all)

  for (;;) {
    // ---- refill idle lanes from the private stack ----
    for (int tries = 0; tries < a.refill_tries; ++tries) {
      const uint64_t need = ballot(!active);
      if (need == 0 || top == bot) break;
      const long long tidx = top - 1;
      if (staged != tidx) {
        if (staged >= bot && staged < top && lane == 0)
          stk_node(a, wid, staged)->next_slot = slot_range(cur_slot, cur_end);
        wave_mem_sync();
        stage_node(stg, (const uint32_t *)stk_node(a, wid, tidx), ptb, lane);
        staged = tidx;
        stg_checked = false;
        const uint32_t sw = (uint32_t)uni((int)stg[15]);
        cur_slot = (int)(sw & 0xFFFFu);
        cur_end = (int)(sw >> 16);
        if (cur_end == 0) cur_end = pt_word(ptb, uni((int)stg[12]), 9);
      }
      if (NARROW && !stg_checked) {
        // a node whose head or written cells lie past the lane's LDS tape goes, with its slot
        // cursor, to the list of the launch with the whole window (tested once per staging)
        const bool past = lane < E13_TW - tw && stg[E13_NODE_W_TAPE + tw + lane] != 0u;
        if (uni((int)stg[0]) > c.qlim || ballot(past) != 0) {
          uint32_t idx = 0;
          uint32_t cap = 0;
          E13Node *wl = nullptr;
          if (lane == 0) {
            idx = atomicAdd(&a.sch->wide_n, 1u);
            cap = a.sch->wide_cap;
            wl = a.sch->wide;
          }
          idx = (uint32_t)uni((int)idx);
          cap = (uint32_t)uni((int)cap);
          if (idx >= cap) { // full: this launch stops cleanly, the host runs the list
            if (lane == 0) atomicMax(&a.sch->halt, 1);
            halt_self = true;
            break;
          }
          uint32_t *o = (uint32_t *)((E13Node *)uni64((long long)wl) + idx);
          const uint32_t v0 = lane == E13_NODE_W_SLOTS ? (uint32_t)slot_range(cur_slot, cur_end) : stg[lane];
          o[lane] = v0;
          if (lane < E13_NODE_WORDS - 64) o[64 + lane] = stg[64 + lane];
          top -= 1; // the node leaves this wave's stack
          staged = -1;
          continue;
        }
        stg_checked = true;
      }
      const int sprob = uni((int)stg[12]);
      const int cnt = __popcll(need);
      const int give = min(cnt, cur_end - cur_slot);
      if (!active) {
        const int r = (int)rank_in(need);
        if (r < give) {
          // child slot -> the value bound to the pending cell (explore(),
          // e13search.c:236-251): slot 0 keeps the base value, slot 1+v binds
          // move v; stage words: 9 budget, 10 symmetric states in use, 11 pending cell
          const int slot = cur_slot + r;
          const int pend = uni((int)stg[11]);
          const int nb0 = uni((int)stg[9]), ni0 = uni((int)stg[10]);
          int nb = nb0, nin2 = ni0, val = -1, ok;
          if (pend < 0) { // a machine paused at a cap: the node is its own single child
            ok = slot == 0;
          } else {
            const int orig = pt_base(ptb, sprob, pend); // move, -1 halt, -2 free (FORCED)
            if (slot == 0) {
              ok = orig >= 0;
              val = orig;
              if (orig == -1) { // keeping a halt: one leaf at the node's stage, 0 steps
                atomicAdd(&pc[sprob * E13_PCNT + 0], 1ull);
                atomicAdd(&pc[sprob * E13_PCNT + 3 + (int)stg[4]], 1ull);
                const uint32_t sd = stg[E13_NODE_W_SEED];
                if (sd < a.nseed_cnt) atomicAdd((unsigned long long *)&a.seed_cnt[2 * sd], 1ull);
              }
            } else {
              const int v = slot - 1;
              const int n2 = sym_next(v, ni0, pt_word(ptb, sprob, 10), pt_word(ptb, sprob, 11));
              ok = (orig == -2 || nb0 > 0) && v != orig && n2 >= 0;
              nb = nb0 - (orig != -2);
              nin2 = n2;
              val = v;
            }
          }
          if (ok) {
            seed = stg[E13_NODE_W_SEED];
            // the machine state, words 0..23 of the staged node
            const uint4 s0 = *(const uint4 *)(stg + 0), s1 = *(const uint4 *)(stg + 4);
            const uint4 s2 = *(const uint4 *)(stg + 8), s3 = *(const uint4 *)(stg + 12);
            const uint4 s4 = *(const uint4 *)(stg + 16), s5 = *(const uint4 *)(stg + 20);
            L.q = (int)s0.x;
            L.row = c.rbl + (int)s0.y * 256;
            L.t = (int)s0.z;
            L.nt = (int)s0.w;
            L.stage = (int)s1.x;
            L.st0 = (int)s1.y;
            L.q0 = (int)s1.z;
            L.flags = (int)s1.w;
            L.ones = (int)s2.x;
            // the block hash only passes through the count gate's machines (always 0
            // there): a constant, so it holds no registers across the loop
            L.hb = GATE == GATE_HASH ? (uint64_t)s4.x | ((uint64_t)s4.y << 32) : 0ull;
            L.ch0 = (uint64_t)s4.z | ((uint64_t)s4.w << 32);
            L.ch1 = (uint64_t)s5.x | ((uint64_t)s5.y << 32);
            L.ch2 = (uint64_t)s5.z | ((uint64_t)s5.w << 32);
            L.flags |= F_DM; // not known: the first step tests the gate
            arm_stage(L, c);
            t0 = L.t;
            budget = nb;
            meta = meta_pack(sprob, pt_word(ptb, sprob, 8), nin2, (int)s3.y + (pend >= 0), (int)s3.z + (pend >= 0));
            // rows: as stored (A), or with edits used up (B: unread cells keep
            // their value), then the pending cell's entry; every stage word is
            // read before any lane column is written (the compiler cannot
            // tell the stage from the columns and would serialise them)
            const uint4 *rsrc = (const uint4 *)(stg + ((nb == 0 && nb0 > 0) ? 92 : 76));
            uint4 rw[4];
#pragma unroll
            for (int k4 = 0; k4 < 4; ++k4) rw[k4] = rsrc[k4];
            uint32_t pr = 0;
            if (pend >= 0) pr = ((const uint32_t *)rsrc)[pend >> 1];
            // the column base as an opaque value: known to be a multiple of 4
            // under 256 apart from its wave part, the compiler would OR the row
            // offsets into it and keep 16 addresses live across the loop
            // instead of folding them into the stores' offsets
            int rb = c.rbl;
            asm volatile("" : "+v"(rb));
#pragma unroll
            for (int k4 = 0; k4 < 4; ++k4) {
              lds_st(c.sb, rb + (4 * k4 + 0) * 256, rw[k4].x);
              lds_st(c.sb, rb + (4 * k4 + 1) * 256, rw[k4].y);
              lds_st(c.sb, rb + (4 * k4 + 2) * 256, rw[k4].z);
              lds_st(c.sb, rb + (4 * k4 + 3) * 256, rw[k4].w);
            }
            if (pend >= 0) { // val is a move here (a base value or one of 0..4nq-1)
              const uint32_t sh = ((uint32_t)pend & 1u) << 4;
              const uint32_t e = enc_move(val, (uint32_t)pend & 1u);
              lds_st(c.sb, rb + (pend >> 1) * 256, (pr & ~(0xFFFFu << sh)) | (e << sh));
            }
            // table codes: as stored (words 24..31) or with edits used up, then the pending cell
            if (c.cbl >= 0) {
              const uint4 *csrc = (const uint4 *)(stg + ((nb == 0 && nb0 > 0) ? STG_W_CODESB : E13_NODE_W_TAB));
              const uint4 c0 = csrc[0], c1 = csrc[1];
              lds_st(c.sb, c.cbl + 0 * 256, c0.x);
              lds_st(c.sb, c.cbl + 1 * 256, c0.y);
              lds_st(c.sb, c.cbl + 2 * 256, c0.z);
              lds_st(c.sb, c.cbl + 3 * 256, c0.w);
              lds_st(c.sb, c.cbl + 4 * 256, c1.x);
              lds_st(c.sb, c.cbl + 5 * 256, c1.y);
              lds_st(c.sb, c.cbl + 6 * 256, c1.z);
              lds_st(c.sb, c.cbl + 7 * 256, c1.w);
              if (pend >= 0)
                *(__attribute__((address_space(3))) uint8_t *)(size_t)(uint32_t)(c.cbl + (pend >> 2) * 256 +
                                                                                  (pend & 3)) = (uint8_t)val;
            }
            // tape: words 32..74 of the stage (word 75, the seed, comes along), the
            // first tw of them (a node that fits the lane's tape is blank past them)
            uint32_t tv[44];
#pragma unroll
            for (int k4 = 0; k4 < 11; ++k4) {
              const uint4 v = *(const uint4 *)(stg + E13_NODE_W_TAPE + 4 * k4);
              tv[4 * k4 + 0] = v.x;
              tv[4 * k4 + 1] = v.y;
              tv[4 * k4 + 2] = v.z;
              tv[4 * k4 + 3] = v.w;
            }
            const LaneTape lt = lane_tape(c);
#pragma unroll
            for (int w = 0; w < E13_TW_MIN; ++w) lds_st(c.sb, lt.tb + w * 256, tv[w]);
            // words E13_TW_MIN .. tw-1: nested tests (structured, so they stay
            // uniform branches, and only as many as the tape has words)
            tape_tail<E13_TW_MIN>(c, lt.tb, tv, lt.tw);
            active = 1;
          }
        }
      }
      cur_slot += give;
      mach += (uint64_t)give;
      if (cur_slot >= cur_end) { // every child slot handed out: pop
        top -= 1;
        staged = -1;
      }
    }
    tn = (CLOCKS ? stamp_if(a.phase_cycles) : 0ull);
    cyc0 += tn - ts;
    ts = tn;
    // ---- this wave's mailbox: a donor may have delivered a node (time: pool/idle) ----
    if (requested && --mail_in <= 0) {
      mail_in = a.mail_every;
      uint64_t f = 0;
      if (lane == 0) f = ld_agent64(&a.mail[wid].flag);
      f = (uint64_t)uni64((long long)f);
      if (f == E13_MAIL_DONE) break; // the search is over (or aborted)
      if (f != last_tag) {
        if (top - bot + 1 > a.stk_cap) {
          if (lane == 0) atomicMax(&a.sch->error, 1);
          broadcast_done(a.mail, a.total_waves, lane);
          break;
        }
        // read until every element carries this delivery's tag
        const uint32_t tag = (uint32_t)f;
        const uint64_t *src = a.mail[wid].w;
        uint64_t e0 = 0, e1 = 0;
        uint32_t spins = 0;
        for (;;) {
          e0 = ld_agent64(src + lane);
          e1 = lane < E13_NODE_WORDS - 64 ? ld_agent64(src + 64 + lane) : ((uint64_t)tag << 32);
          const bool ok = (uint32_t)(e0 >> 32) == tag && (uint32_t)(e1 >> 32) == tag;
          if (ballot(!ok) == 0) break;
          if (++spins > a.spin_limit) break;
          __builtin_amdgcn_s_sleep(2);
        }
        if (spins > a.spin_limit) {
          if (lane == 0) atomicMax(&a.sch->error, 2);
          broadcast_done(a.mail, a.total_waves, lane);
          break;
        }
        uint32_t *dst = (uint32_t *)stk_node(a, wid, top);
        dst[lane] = (uint32_t)e0;
        if (lane < E13_NODE_WORDS - 64) dst[64 + lane] = (uint32_t)e1;
        // the donor counted the delivery as busy: it becomes this wave's count,
        // or is returned when this wave was counted already
        if (busy && lane == 0) atomicSub(&a.sch->busy, 1);
        last_tag = f;
        requested = 0;
        busy = 1;
        idle_since = 0;
        top += 1;
        ++w_takes;
        TRACE_EV(TR_TAKE);
        continue; // refill from it at once
      }
    }
    // ---- an empty private stack and idle lanes: ask for work (one bit in the hungry map) ----
    // (not while its only running lane is a long runner: that machine keeps the scalar loop)
    if (!requested && top == bot) {
      const int idle = __popcll(ballot(!active));
      const bool hold = a.evict_after > 0 && idle == 63 && ballot(active && L.t >= a.evict_after) != 0;
      if ((idle >= a.pool_min_idle || idle == 64) && !hold) {
        if (lane == 0) atomicOr((unsigned long long *)&a.hungry[wid >> 6], 1ull << (wid & 63));
        requested = 1;
        mail_in = 1;
      }
    }
    const uint64_t act = ballot(active);
    if (act == 0) {
      if (NARROW && halt_self) { // nothing runs: keep the staged node's cursor and leave
        if (staged >= bot && staged < top && lane == 0)
          stk_node(a, wid, staged)->next_slot = slot_range(cur_slot, cur_end);
        break;
      }
      if (top != bot) continue;
      // no work here: leave the busy count; the wave that takes it to zero ends the search
      if (busy) {
        vm_drain(); // every delivery of this wave is complete before it stops counting as busy
        int last = 0;
        if (lane == 0) last = atomicSub(&a.sch->busy, 1) == 1;
        busy = 0;
        TRACE_EV(TR_IDLE);
        if (uni(last)) {
          broadcast_done(a.mail, a.total_waves, lane);
          break;
        }
      }
      const uint64_t now = __builtin_readsteadycounter();
      if (idle_since == 0) idle_since = now;
      if ((++idle_polls & 63) == 0) { // an abort may have overwritten the end mark of this mailbox
        int err = 0, stp = 0;
        if (lane == 0) err = ld_agent_i(&a.sch->error);
        if (a.stop != nullptr && lane == 0)
          stp = (int)__hip_atomic_load(a.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        if (NARROW && lane == 0) stp |= ld_agent_i(&a.sch->halt);
        if (uni(err) || uni(stp)) break; // a clean stop: nothing here to keep (a delivery still on its way
                                         // stays in the mailbox for the host)
      }
      if (now - idle_since > (uint64_t)a.idle_limit) { // watchdog: no delivery and no end for this long
        if (lane == 0) atomicMax(&a.sch->error, 3);
        broadcast_done(a.mail, a.total_waves, lane);
        break;
      }
      for (int z = 0; z < a.idle_naps; ++z) __builtin_amdgcn_s_sleep(8);
      tn = (CLOCKS ? stamp_if(a.phase_cycles) : 0ull);
      cyc3 += tn - ts;
      ts = tn;
      continue;
    }
    idle_since = 0;
    w_rounds++;
    w_slots += (uint64_t)__popcll(act);
    if ((w_rounds & 63u) == 0u) TRACE_EV2(TR_ROUND, __popcll(act) | (top - bot > 2 ? 0x80 : 0));
    // shared state, read every few rounds (staggered over the waves): the abort
    // flag and the hungry map; the loads complete while the lanes step
    const bool refresh = --refresh_in <= 0;
    int s_err = 0, s_stop = 0;
    uint64_t hw = 0; // hungry-map word `lane` (maps of up to 4096 waves)
    if (refresh) {
      if (lane == 0) s_err = ld_agent_i(&a.sch->error);
      if (a.stop != nullptr && lane == 0)
        s_stop = (int)__hip_atomic_load(a.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      if (NARROW && lane == 0) s_stop |= ld_agent_i(&a.sch->halt);
      const bool can_give = top - bot >= 2 || (top - bot == 1 && staged == top - 1 && cur_end - cur_slot >= 2) ||
                            (a.evict_after > 0 && top == bot && __popcll(act) >= 2);
      if (lane < a.hungry_words && can_give) hw = ld_agent64(&a.hungry[lane]);
    }
    tn = (CLOCKS ? stamp_if(a.phase_cycles) : 0ull);
    cyc3 += tn - ts;
    ts = tn;
    // ---- a round of steps (time: stepping) ----
    // a lone running lane with nothing to hand out takes the scalar loop for a longer round
    int ev = EV_NONE, evcell = -1;
    const bool solo = GATE == GATE_COUNT && a.solo > 0 && __popcll(act) == 1 && top == bot;
    if (solo) { // the whole wave runs the lone lane's loop
      solo_memo_reset(active, c, act, mach, solo_last);
      ev = run_solo(L, c, a.solo, evcell, __ffsll((long long)act) - 1);
      if (!active) ev = EV_NONE;
    } else if (active) {
      // with nodes to hand out, the round may end early once most lanes have stopped
      const bool cut = a.cut_live > 1 && top > bot;
      if (a.skip_settled) {
        if (cut) ev = run_slice<GATE, true, true>(L, c, a.slice, a.cut_live, evcell);
        else ev = run_slice<GATE, false, true>(L, c, a.slice, 0, evcell);
      } else {
        if (cut) ev = run_slice<GATE, true, false>(L, c, a.slice, a.cut_live, evcell);
        else ev = run_slice<GATE, false, false>(L, c, a.slice, 0, evcell);
      }
    }
    tn = (CLOCKS ? stamp_if(a.phase_cycles) : 0ull);
    cyc1 += tn - ts;
    ts = tn;
    // ---- a clean stop: every running machine is paused as a resumable node, as at
    // a launch cap, and goes onto this wave's stack below ----
    const bool stopping = (refresh && uni(s_stop) != 0) || (NARROW && halt_self);
    if (stopping && active && ev == EV_NONE) ev = EV_CAP;
    // ---- finished machines ----
    // ---- long runners to hungry waves: with nothing on its stack to give, a wave
    // pauses machines that are far along (evict_after steps from the blank tape:
    // past the early stage bounds they tend to run to the late deadlines) as
    // resumable nodes, as at a launch cap, so the donation below hands them out
    // and each runs alone, in the scalar loop ----
    if (refresh && a.evict_after > 0 && top == bot) {
      const uint64_t running = ballot(active && ev == EV_NONE);
      if (__popcll(running) >= 2 && ballot(hw != 0) != 0) {
        const uint64_t cand = ballot(active && ev == EV_NONE && L.t >= a.evict_after);
        const int n_ev = min(min(__popcll(cand), a.donate_every), __popcll(running) - 1);
        if (active && ev == EV_NONE && ((cand >> lane) & 1ull) && (int)rank_in(cand) < n_ev) ev = EV_CAP;
      }
    }
    const bool fin = active && ev != EV_NONE;
    const int prob = meta_prob(meta), nq = meta_nq(meta);
    if (fin) {
      const unsigned long long d = (unsigned long long)(L.t - t0);
      w_steps += (uint32_t)d;
      if (ev == EV_DEAD) {
        atomicAdd(&pc[prob * E13_PCNT + 0], 1ull);
        atomicAdd(&pc[prob * E13_PCNT + 3 + L.stage], 1ull);
      } else if (ev == EV_SURV) {
        atomicAdd(&pc[prob * E13_PCNT + 0], 1ull);
        atomicAdd(&pc[prob * E13_PCNT + 2], 1ull);
        const uint32_t sidx = atomicAdd(&a.ctr->nsurv, 1u);
        if (sidx < a.surv_cap) {
          E13SurvRec *o = a.surv + sidx;
          uint32_t *tw = (uint32_t *)o->tab;
#pragma unroll
          for (int i = 0; i < E13_MAXCELLS / 4; ++i) tw[i] = codes_word(c, nq, i);
          o->st0 = L.st0;
          o->head0 = L.q0 - E13_QOFF;
          o->abs_ok = L.flags & 1;
          o->rel_ok = (L.flags >> 1) & 1;
          o->steps = L.t;
          o->prob = prob;
          o->ch[0] = L.ch0;
          o->ch[1] = L.ch1;
          o->ch[2] = L.ch2;
        }
      }
      atomicAdd(&pc[prob * E13_PCNT + 1], d);
      if (seed < a.nseed_cnt) { // cost logging: this subtree's leaves and steps
        if (ev == EV_DEAD || ev == EV_SURV) atomicAdd((unsigned long long *)&a.seed_cnt[2 * seed], 1ull);
        atomicAdd((unsigned long long *)&a.seed_cnt[2 * seed + 1], d);
      }
    }
    if (ballot(w_steps >= 0x80000000u)) { // fold before a lane's 32-bit count could wrap (each d is below 2^31)
      w_steps_wave += wave_sum64(w_steps);
      w_steps = 0;
    }
    // ---- live children onto the private stack ----
    const bool emit = fin && (ev == EV_STOP || ev == EV_CAP);
    const uint64_t em = ballot(emit);
    if (em) {
      const int n = __popcll(em);
      if (top - bot + n > a.stk_cap) {
        if (lane == 0) atomicMax(&a.sch->error, 1);
        broadcast_done(a.mail, a.total_waves, lane);
        break;
      }
      if (staged >= bot && staged < top && lane == 0)
        stk_node(a, wid, staged)->next_slot = slot_range(cur_slot, cur_end);
      if (emit) {
        const int r = (int)rank_in(em);
        const int nintro = meta_nintro(meta), depth = meta_depth(meta), nbound = meta_nbound(meta);
        write_node(c, stk_node(a, wid, top + r), L, budget, nintro, ev == EV_STOP ? evcell : -1, prob, depth, nbound,
                   nq, seed);
        if (r == n - 1) stage_head(c, stg, L, budget, nintro, ev == EV_STOP ? evcell : -1, prob, depth, nbound);
      }
      stage_body(c, stg, 63 - __builtin_clzll(em), seed, nq, lane);
      stage_rows(stg, ptb, lane);
      top += n;
      staged = top - 1;
      stg_checked = false; // a child paused past the lane tape may be on top
      cur_slot = 0;
      cur_end = pt_word(ptb, uni((int)stg[12]), 9);
    }
    if (fin) active = 0;
    if (requested && top > bot) { // work of its own again: withdraw the request unless a donor took it
      const uint64_t bit = 1ull << (wid & 63);
      int withdrawn = 0;
      if (lane == 0) withdrawn = (atomicAnd((unsigned long long *)&a.hungry[wid >> 6], ~bit) & bit) != 0;
      if (uni(withdrawn)) requested = 0; // else a delivery is on its way: keep polling the mailbox
    }
    tn = (CLOCKS ? stamp_if(a.phase_cycles) : 0ull);
    cyc2 += tn - ts;
    ts = tn;
    if (stopping) { // the staged node's cursor goes with it; the stack is this wave's whole state
      if (staged >= bot && staged < top && lane == 0)
        stk_node(a, wid, staged)->next_slot = slot_range(cur_slot, cur_end);
      break;
    }
    if (refresh) {
      if (uni(s_err)) break;
      // ---- deliver work to hungry waves: the oldest node, or with a single
      // node, the upper part of the staged node's child slots ----
      uint64_t nz = ballot(hw != 0);
      int given = 0;
      const int idle_now = __popcll(ballot(!active));
      for (;;) {
        // a whole node: keep one unless lanes still run (they keep the wave busy)
        const bool whole = top - bot >= 2 || (top - bot == 1 && ballot(active) != 0 && a.evict_after > 0);
        const int rem = cur_end - cur_slot, keep = max(idle_now, rem / 2);
        const bool split = !whole && top - bot == 1 && staged == top - 1 && rem >= 2 && keep < rem;
        if (nz == 0 || given >= a.donate_every || !(whole || split)) break;
        // a word and a bit chosen by this wave's id, so donors spread over the hungry waves
        const int rot = (wid * 7) & 63;
        const uint64_t nzr = (nz >> rot) | (rot ? (nz << (64 - rot)) : 0ull);
        const int wsel = (__ffsll((long long)nzr) - 1 + rot) & 63;
        const uint64_t word = (uint64_t)__shfl((long long)hw, wsel);
        const int brot = (wid * 13 + given) & 63;
        const uint64_t wr = (word >> brot) | (brot ? (word << (64 - brot)) : 0ull);
        const int bsel = (__ffsll((long long)wr) - 1 + brot) & 63;
        const uint64_t bit = 1ull << bsel;
        int won = 0;
        uint64_t dseq = 0; // the receiver's delivery number
        const int dst_w = wsel * 64 + bsel;
        if (lane == 0) {
          won = (atomicAnd((unsigned long long *)&a.hungry[wsel], ~bit) & bit) != 0;
          if (won) {
            atomicAdd(&a.sch->busy, 1); // the delivery counts as busy until it is taken
            dseq = atomicAdd((unsigned long long *)&a.mail[dst_w].seq, 1ull);
          }
        }
        won = uni(won);
        if (lane == wsel) hw &= ~bit;
        nz = ballot(hw != 0);
        if (!won) continue;
        uint32_t slots_word = 0xFFFFFFFFu; // the donated copy's slot range (split only)
        if (whole && staged == bot) { // a delivery may sit above it: its slot cursor travels with it
          if (lane == 0) stk_node(a, wid, staged)->next_slot = slot_range(cur_slot, cur_end);
          staged = -1;
        }
        if (!whole) { // keep [cur_slot, cur_slot + keep), give the rest
          slots_word = slot_range(cur_slot + keep, cur_end);
          cur_end = cur_slot + keep;
        }
        wave_mem_sync();
        const uint32_t *src = (const uint32_t *)stk_node(a, wid, whole ? bot : staged);
        uint64_t *dst = a.mail[dst_w].w;
        const uint64_t tag = (uint64_t)uni64((long long)dseq) % 0xFFFFFFFFull + 1ull; // 1 .. 2^32 - 1
        uint32_t v0 = src[lane];
        const uint32_t v1 = lane < E13_NODE_WORDS - 64 ? src[64 + lane] : 0u;
        if (!whole && lane == E13_NODE_W_SLOTS) v0 = slots_word;
        st_agent64(dst + lane, (tag << 32) | v0);
        if (lane < E13_NODE_WORDS - 64) st_agent64(dst + 64 + lane, (tag << 32) | v1);
        vm_drain(); // the elements are issued and complete before the flag
        if (lane == 0) st_agent64(&a.mail[dst_w].flag, tag);
        if (whole) bot += 1;
        ++given;
        ++w_gives;
        TRACE_EV(TR_GIVE);
      }
      refresh_in = given > 0 || nz != 0 ? 2 : (int)a.sched_every; // look again soon while waves wait
    }
    tn = (CLOCKS ? stamp_if(a.phase_cycles) : 0ull);
    cyc3 += tn - ts;
    ts = tn;
  }
  // ---- flush ----
  if (a.wave_state != nullptr && lane == 0) {
    a.wave_state[3 * wid + 0] = bot;
    a.wave_state[3 * wid + 1] = top;
    a.wave_state[3 * wid + 2] = (int64_t)last_tag;
  }
  if (lane == 0) {
    atomicAdd(&a.sch->exited, 1);
    atomicAdd(&a.sch->donations, (int)w_gives);
    atomicAdd(&a.sch->pool_takes, (int)w_takes);
  }
  TRACE_EV(TR_EXIT);
  if (CLOCKS && a.trace != nullptr && lane == 0) a.trace[(size_t)wid * (size_t)a.trace_cap] = tr_n;
  cyc3 += (CLOCKS ? stamp_if(a.phase_cycles) : 0ull) - ts;
  if (lane == 0) {
    atomicAdd((unsigned long long *)&a.ctr->cyc[0], (unsigned long long)cyc0);
    atomicAdd((unsigned long long *)&a.ctr->cyc[1], (unsigned long long)cyc1);
    atomicAdd((unsigned long long *)&a.ctr->cyc[2], (unsigned long long)cyc2);
    atomicAdd((unsigned long long *)&a.ctr->cyc[3], (unsigned long long)cyc3);
    atomicAdd((unsigned long long *)&a.ctr->rounds, (unsigned long long)w_rounds);
    atomicAdd((unsigned long long *)&a.ctr->lane_slots, (unsigned long long)w_slots);
  }
  const uint64_t s_steps = w_steps_wave + wave_sum64(w_steps);
  if (lane == 0) atomicAdd((unsigned long long *)&a.ctr->steps, (unsigned long long)s_steps);
  __syncthreads();
  for (int i = threadIdx.x; i < a.np * E13_PCNT; i += blockDim.x)
    if (pc[i]) atomicAdd((unsigned long long *)&a.pcnt[i], pc[i]);
}

template <int GATE, bool NARROW>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NARROW ? 4 : 2))) k_search(E13SearchArgs a) {
  search_body<GATE, NARROW, !(NARROW && GATE == GATE_COUNT)>(a);
}
// The short-tape count-gate search with the per-phase clock reads and the trace.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_search_clocked(E13SearchArgs a) {
  search_body<GATE_COUNT, true, true>(a);
}

// ---------------------------------------------------------------------------
// Probe estimator (probe(), e13search.c:256-313; estimate mode, :473-510).
// A probe is one random root-to-leaf path: at each node on it every child is
// run to its next stop and counted with the path weight w, then the path goes
// on through a uniformly drawn live child and w grows by the number of live
// children.  A wave keeps PK probes in flight, each in a slot of LDS holding
// the node on its path (tape, state, table), its children's results and the
// probe's sums; idle lanes take child runs from any slot, so the long child of
// one level does not hold the wave.  When a level's children are all in, the
// wave adds them to the sums in child order (slot 0 keeps the base value,
// slot 1+v binds move v: the reference's child order, so its additions and
// their rounding are the reference's), draws the next child from the probe's
// own stream, and one lane re-runs that child from the node to make the next
// node.  Every cell starts unread (the reference's probe starts with fixed[]
// all zero), so the path also stops at the first read of one-child cells.
// ---------------------------------------------------------------------------
namespace {

__device__ __forceinline__ uint64_t probe_mix(uint64_t &s) { // splitmix64, as e13search.c:98-103
  s += 0x9E3779B97F4A7C15ull;
  uint64_t z = s;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Probe p's stream: its own splitmix64 state, started at a hashed point.
__device__ __forceinline__ uint64_t probe_stream(uint64_t seed, uint64_t p) {
  uint64_t s = 0x5EED0000ull + (seed << 40) + p;
  return probe_mix(s);
}

#define PK 6          // probes in flight per wave
// a slot, in words from its base
#define PS_TAPE 0     // the node's tape (43 words)
#define PS_REG 44     // the node's state: q st t nt stage st0 q0 flags ones hb(2) ch0(2) ch1(2) ch2(2)
#define PS_BUDGET 61  //   edits left, symmetric states in use, the cell to bind
#define PS_NINTRO 62
#define PS_PEND 63
#define PS_TAB 64     // the path's table, 32 codes (8 words)
#define PS_DT 72      // steps of each child of the node (64 words)
#define PS_LIST 136   // the children to run, in child order (64 bytes)
#define PS_EXIST 152  // existing children (2 words), ended children (2), survivors (2)
#define PS_ENDED 154
#define PS_SURV 156
#define PS_PH 158     // phase
#define PS_P 159      // probe index
#define PS_NEXT 160   // children handed out
#define PS_COUNT 161  // children to run
#define PS_LEFT 162   // children still running
#define PS_CHOSEN 163 // the drawn child
#define PS_RNG 164    // the probe's stream (2 words)
#define PS_W 166      // the path weight (double)
#define PS_SUM 168    // leaves, nodes, steps, survivors (doubles)
#define PS_WORDS 176
enum { PH_EMPTY = 0, PH_LEVEL = 1, PH_CHILDREN = 2, PH_DRAWN = 3, PH_ADVANCE = 4 };
#define ADV_CHILD 64  // a lane's task is the drawn child's re-run

__device__ __forceinline__ uint64_t lds_ld64(const Ctx &c, int off) {
  return (uint64_t)lds_ld(c.sb, off) | ((uint64_t)lds_ld(c.sb, off + 4) << 32);
}
__device__ __forceinline__ void lds_st64(const Ctx &c, int off, uint64_t v) {
  lds_st(c.sb, off, (uint32_t)v);
  lds_st(c.sb, off + 4, (uint32_t)(v >> 32));
}
__device__ __forceinline__ uint64_t uni64(uint64_t v) {
  return (uint64_t)(uint32_t)uni((int)(uint32_t)v) | ((uint64_t)(uint32_t)uni((int)(uint32_t)(v >> 32)) << 32);
}

// The children of one node in child order (e13search.c:289-300): nodes += w,
// steps += w * dt, leaves += w for an ended child, survivors += w for a
// survivor.  No contraction: every product and sum rounds as the reference's.
__device__ __forceinline__ void probe_level_add(const Ctx &c, int sb, uint64_t exist, uint64_t ended,
                                                uint64_t surv) {
#pragma clang fp contract(off)
  const double w = __builtin_bit_cast(double, lds_ld64(c, sb + PS_W * 4));
  double leaves = __builtin_bit_cast(double, lds_ld64(c, sb + (PS_SUM + 0) * 4));
  double nodes = __builtin_bit_cast(double, lds_ld64(c, sb + (PS_SUM + 2) * 4));
  double steps = __builtin_bit_cast(double, lds_ld64(c, sb + (PS_SUM + 4) * 4));
  double sv = __builtin_bit_cast(double, lds_ld64(c, sb + (PS_SUM + 6) * 4));
  while (exist) {
    const int s = __builtin_ctzll(exist);
    exist &= exist - 1;
    const int d = (int)lds_ld(c.sb, sb + (PS_DT + s) * 4);
    nodes = nodes + w;
    steps = steps + w * (double)d;
    if ((ended >> s) & 1ull) {
      leaves = leaves + w;
      if ((surv >> s) & 1ull) sv = sv + w;
    }
  }
  if ((threadIdx.x & 63u) == 0) {
    lds_st64(c, sb + (PS_SUM + 0) * 4, __builtin_bit_cast(uint64_t, leaves));
    lds_st64(c, sb + (PS_SUM + 2) * 4, __builtin_bit_cast(uint64_t, nodes));
    lds_st64(c, sb + (PS_SUM + 4) * 4, __builtin_bit_cast(uint64_t, steps));
    lds_st64(c, sb + (PS_SUM + 6) * 4, __builtin_bit_cast(uint64_t, sv));
  }
}

// Child ch of a node whose pending cell has base code orig (e13search.c:278-287):
// 1 if it exists, with its value, edits left and symmetric states in use.
__device__ __forceinline__ int probe_child(int ch, int orig, int budget, int nintro, const E13Problem *pb, int &val,
                                           int &nb, int &ni) {
  val = -1;
  nb = budget;
  ni = nintro;
  if (ch == 0) {
    if (orig == -2) return 0;
    val = orig;
    return 1;
  }
  const int v = ch - 1;
  if (!(orig == -2 || budget > 0) || v == orig) return 0;
  const int n2 = sym_next(v, nintro, pb->sym_lo, pb->sym_hi);
  if (n2 < 0) return 0;
  val = v;
  ni = n2;
  if (orig != -2) nb = budget - 1;
  return 1;
}

// The sums of a level whose only child is child ch (nodes += w, steps += w *
// dt; leaves and survivors when it ended), written back; e13search.c:289-300.
__device__ __forceinline__ void probe_single_add(const Ctx &c, int sb, int dt, bool ended, bool surv) {
#pragma clang fp contract(off)
  const double w = __builtin_bit_cast(double, lds_ld64(c, sb + PS_W * 4));
  const double nodes = __builtin_bit_cast(double, lds_ld64(c, sb + (PS_SUM + 2) * 4)) + w;
  const double steps = __builtin_bit_cast(double, lds_ld64(c, sb + (PS_SUM + 4) * 4)) + w * (double)dt;
  lds_st64(c, sb + (PS_SUM + 2) * 4, __builtin_bit_cast(uint64_t, nodes));
  lds_st64(c, sb + (PS_SUM + 4) * 4, __builtin_bit_cast(uint64_t, steps));
  if (ended) {
    const double leaves = __builtin_bit_cast(double, lds_ld64(c, sb + (PS_SUM + 0) * 4)) + w;
    lds_st64(c, sb + (PS_SUM + 0) * 4, __builtin_bit_cast(uint64_t, leaves));
    if (surv) {
      const double sv = __builtin_bit_cast(double, lds_ld64(c, sb + (PS_SUM + 6) * 4)) + w;
      lds_st64(c, sb + (PS_SUM + 6) * 4, __builtin_bit_cast(uint64_t, sv));
    }
  }
}

// The lane's machine (the path's child that was bound to tval at the pending
// cell) stopped before reading cell evcell: it is the node of the probe's next
// level.  When that level's only child runs on from it (the cell keeps a move:
// no edits left and not a free cell), the lane binds the cell in its own rows
// and goes on as that child (returns true); otherwise it writes the node to the
// slot for the wave to expand.
__device__ __forceinline__ bool probe_next_node(const Ctx &c, const Lane &L, uint8_t *sbytes, int sb, int evcell,
                                                int &tval, int tnb, int tni, const int8_t *base) {
  const int pend = (int)lds_ld(c.sb, sb + PS_PEND * 4);
  sbytes[sb + PS_TAB * 4 + pend] = (uint8_t)tval;
  lds_st(c.sb, sb + PS_PEND * 4, (uint32_t)evcell);
  lds_st(c.sb, sb + PS_BUDGET * 4, (uint32_t)tnb);
  lds_st(c.sb, sb + PS_NINTRO * 4, (uint32_t)tni);
  const int orig = base[evcell];
  if (orig >= 0 && tnb == 0) { // one child, the keep child (e13search.c:279-282)
    const int ra = c.rbl + (evcell >> 1) * 256;
    const uint32_t e = enc_entry(orig, (uint32_t)(evcell & 1));
    const uint32_t row = lds_ld(c.sb, ra);
    lds_st(c.sb, ra, (evcell & 1) ? ((row & 0xFFFFu) | (e << 16)) : ((row & 0xFFFF0000u) | e));
    tval = orig;
    lds_st64(c, sb + PS_EXIST * 4, 1ull);
    lds_st(c.sb, sb + PS_NEXT * 4, 1u);
    lds_st(c.sb, sb + PS_COUNT * 4, 1u);
    lds_st(c.sb, sb + PS_LEFT * 4, 1u);
    lds_st(c.sb, sb + PS_PH * 4, PH_CHILDREN);
    return true;
  }
  for (int wd = 0; wd < E13_TW; ++wd) lds_st(c.sb, sb + (PS_TAPE + wd) * 4, lds_ld(c.sb, c.tbl + wd * 256));
  const int r0 = sb + PS_REG * 4;
  lds_st(c.sb, r0, (uint32_t)L.q);
  lds_st(c.sb, r0 + 4, (uint32_t)((L.row - c.rbl) >> 8));
  lds_st(c.sb, r0 + 8, (uint32_t)L.t);
  lds_st(c.sb, r0 + 12, (uint32_t)L.nt);
  lds_st(c.sb, r0 + 16, (uint32_t)L.stage);
  lds_st(c.sb, r0 + 20, (uint32_t)L.st0);
  lds_st(c.sb, r0 + 24, (uint32_t)L.q0);
  lds_st(c.sb, r0 + 28, (uint32_t)L.flags);
  lds_st(c.sb, r0 + 32, (uint32_t)L.ones);
  lds_st64(c, r0 + 36, L.hb);
  lds_st64(c, r0 + 44, L.ch0);
  lds_st64(c, r0 + 52, L.ch1);
  lds_st64(c, r0 + 60, L.ch2);
  lds_st(c.sb, sb + PS_PH * 4, PH_LEVEL);
  return false;
}

} // namespace

template <int GATE>
__global__ void __launch_bounds__(256) k_probe(E13ProbeArgs a) {
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
  c.dlb = -1;
  c.cbl = -1;
  c.tw = E13_TW;
  c.qlim = E13_QMAX;
  c.mvb = c.tbl - lane * 4 + SOLO_MEMO_ROWS * 256; // memo values in tape words 16..31
  if (lds_base(smem) != 0u) { // the absolute LDS addressing needs the dynamic area at address 0
    if (threadIdx.x == 0) atomicOr(&a.ctr->overflow, 4u);
    return;
  }
  if (GATE == GATE_HASH) fill_keyb(c.sb, c.keyb, a.key);
  // this wave's PK slots, after the tapes, the block keys and 64 bytes per wave
  const int slots = c.keyb + (GATE == GATE_HASH ? KEYB_WORDS * 8 : 0) + nwaves * 64 + wave * (PK * PS_WORDS * 4);
  for (int i = lane; i < PK * PS_WORDS; i += 64) lds_st(c.sb, slots + i * 4, 0u); // every slot empty
  c.ckb = c.keyb + (GATE == GATE_HASH ? KEYB_WORDS * 8 : 0) + nwaves * 64 + nwaves * (PK * PS_WORDS * 4);
  fill_ckeys(c.ckb, a.key);
  __syncthreads();
  uint8_t *sbytes = (uint8_t *)smem;
  const int nq = a.nq, nslot = 4 * nq + 1;
  const int8_t *base = a.pb->base;
  Lane L;
  L.dlcap = -1;
  L.tcap = 0x7FFFFFFF;
  int active = 0, tslot = 0, tchild = 0, t0 = 0, tval = 0, tnb = 0, tni = 0;
  bool exhausted = false;
  uint64_t my_steps = 0, adv_steps = 0, w_slots = 0, w_rounds = 0, fills = 0, solo_last = ~0ull;
  for (;;) {
    bool progress = false;
    int busy = 0;
    // ---- the slots' bookkeeping (wave-uniform) ----
    for (int k = 0; k < PK; ++k) {
      const int sb = slots + k * (PS_WORDS * 4);
      int ph = uni((int)lds_ld(c.sb, sb + PS_PH * 4));
      if (ph == PH_EMPTY && !exhausted) {
        uint32_t p = 0;
        if (lane == 0) p = atomicAdd(&a.ctr->item, 1u);
        p = (uint32_t)uni((int)p);
        const E13Node *rt = a.roots + (size_t)((a.p0 + p) / a.per_root);
        const int rpend = p < a.n ? uni((int)rt->pend) : 0;
        if (p >= a.n) {
          exhausted = true;
        } else if (rpend < 0 || rpend >= 2 * nq) { // a machine paused at a cap has no pending cell: no probe
          if (lane < 4) a.out[4 * (size_t)p + lane] = -1.0;
          progress = true;
        } else {
          // the probe's first node: root (a.p0 + p) / per_root, a stopped
          // machine with its pending cell (the estimate's root is the blank
          // tape stopped before the read of cell 0 at t = 0, e13search.c:268-271)
          const uint32_t *nw = (const uint32_t *)rt;
          if (lane < E13_TW) lds_st(c.sb, sb + (PS_TAPE + lane) * 4, nw[E13_NODE_W_TAPE + lane]);
          if (lane < 20) { // q st t nt stage st0 q0 flags ones | hb ch0 ch1 ch2 | budget nintro pend
            const int w = lane < 9 ? lane : (lane < 17 ? lane + 7 : lane - 8);
            lds_st(c.sb, sb + (PS_REG + lane) * 4, nw[w]);
          }
          if (lane < E13_MAXCELLS) sbytes[sb + PS_TAB * 4 + lane] = rt->tab[lane];
          if (lane == 0) { // nodes 1, steps 0 (e13search.c:268-271)
            lds_st(c.sb, sb + PS_P * 4, p);
            lds_st64(c, sb + PS_RNG * 4, probe_stream(a.seed, a.p0 + p));
            lds_st64(c, sb + PS_W * 4, __builtin_bit_cast(uint64_t, 1.0));
            lds_st64(c, sb + (PS_SUM + 0) * 4, 0ull);
            lds_st64(c, sb + (PS_SUM + 2) * 4, __builtin_bit_cast(uint64_t, 1.0));
            lds_st64(c, sb + (PS_SUM + 4) * 4, 0ull);
            lds_st64(c, sb + (PS_SUM + 6) * 4, 0ull);
          }
          ph = PH_LEVEL;
          progress = true;
          wave_lds_sync();
        }
      }
      if (ph == PH_LEVEL) { // the node's children, lane = child (e13search.c:278-287)
        const int pend = uni((int)lds_ld(c.sb, sb + PS_PEND * 4));
        const int budget = uni((int)lds_ld(c.sb, sb + PS_BUDGET * 4));
        const int nintro = uni((int)lds_ld(c.sb, sb + PS_NINTRO * 4));
        const int orig = base[pend];
        int val, nb, ni;
        const bool ok = lane < nslot && probe_child(lane, orig, budget, nintro, a.pb, val, nb, ni);
        const bool hk = ok && lane == 0 && orig == -1; // keeping a halt: dies on the read, 0 steps (:164-166)
        const uint64_t exist = ballot(ok), run = ballot(ok && !hk);
        if (ok && !hk) sbytes[sb + PS_LIST * 4 + (int)rank_in(run)] = (uint8_t)lane;
        if (lane == 0) {
          lds_st(c.sb, sb + PS_DT * 4, 0u);
          lds_st64(c, sb + PS_EXIST * 4, exist);
          lds_st64(c, sb + PS_ENDED * 4, hk ? 1ull : 0ull);
          lds_st64(c, sb + PS_SURV * 4, 0ull);
          lds_st(c.sb, sb + PS_NEXT * 4, 0u);
          lds_st(c.sb, sb + PS_COUNT * 4, (uint32_t)__popcll(run));
          lds_st(c.sb, sb + PS_LEFT * 4, (uint32_t)__popcll(run));
        }
        ph = PH_CHILDREN;
        progress = true;
        wave_lds_sync();
      }
      if (ph == PH_CHILDREN) {
        const int next = uni((int)lds_ld(c.sb, sb + PS_NEXT * 4));
        const int count = uni((int)lds_ld(c.sb, sb + PS_COUNT * 4));
        const int left = uni((int)lds_ld(c.sb, sb + PS_LEFT * 4));
        if (next == count && left == 0) { // every child is in
          const uint64_t exist = uni64(lds_ld64(c, sb + PS_EXIST * 4));
          const uint64_t ended = uni64(lds_ld64(c, sb + PS_ENDED * 4));
          const uint64_t surv = uni64(lds_ld64(c, sb + PS_SURV * 4));
          probe_level_add(c, sb, exist, ended, surv);
          wave_lds_sync();
          const uint64_t live = exist & ~ended;
          if (live == 0) { // e13search.c:303
            if (lane < 4) {
              const uint32_t p = lds_ld(c.sb, sb + PS_P * 4);
              a.out[4 * (size_t)p + lane] = __builtin_bit_cast(double, lds_ld64(c, sb + (PS_SUM + 2 * lane) * 4));
            }
            ph = PH_EMPTY;
          } else { // draw the child the path goes on with (e13search.c:304-311)
            uint64_t rng = uni64(lds_ld64(c, sb + PS_RNG * 4));
            const uint32_t nlive = (uint32_t)__popcll(live);
            const uint32_t r = (uint32_t)(probe_mix(rng) % (uint64_t)nlive);
            uint64_t m = live;
            for (uint32_t j = 0; j < r; ++j) m &= m - 1;
            const double w = __builtin_bit_cast(double, lds_ld64(c, sb + PS_W * 4)) * (double)nlive;
            if (lane == 0) {
              lds_st64(c, sb + PS_RNG * 4, rng);
              lds_st64(c, sb + PS_W * 4, __builtin_bit_cast(uint64_t, w));
              lds_st(c.sb, sb + PS_CHOSEN * 4, (uint32_t)__builtin_ctzll(m));
            }
            ph = PH_DRAWN;
          }
          if (lane == 0) lds_st(c.sb, sb + PS_PH * 4, (uint32_t)ph);
          progress = true;
          wave_lds_sync();
        }
      }
      if (lane == 0) lds_st(c.sb, sb + PS_PH * 4, (uint32_t)ph);
      busy += ph != PH_EMPTY;
    }
    wave_lds_sync();
    // ---- work for idle lanes (wave-uniform) ----
    bool fresh = false;
    uint64_t need = ballot(!active);
    for (int k = 0; k < PK && need; ++k) {
      const int sb = slots + k * (PS_WORDS * 4);
      const int ph = uni((int)lds_ld(c.sb, sb + PS_PH * 4));
      if (ph == PH_DRAWN) { // the drawn child's re-run makes the next node
        if (lane == __builtin_ctzll(need)) {
          fresh = true;
          tslot = k;
          tchild = ADV_CHILD;
        }
        if (lane == 0) lds_st(c.sb, sb + PS_PH * 4, PH_ADVANCE);
        progress = true;
      } else if (ph == PH_CHILDREN) {
        const int next = uni((int)lds_ld(c.sb, sb + PS_NEXT * 4));
        const int count = uni((int)lds_ld(c.sb, sb + PS_COUNT * 4));
        const int give = min(__popcll(need), count - next);
        if (give > 0) {
          const int r = (int)rank_in(need);
          if (((need >> lane) & 1ull) && r < give) {
            fresh = true;
            tslot = k;
            tchild = sbytes[sb + PS_LIST * 4 + next + r];
          }
          if (lane == 0) lds_st(c.sb, sb + PS_NEXT * 4, (uint32_t)(next + give));
          progress = true;
        }
      }
      need = ballot(!active && !fresh);
    }
    if (ballot(fresh)) fills++;
    wave_lds_sync();
    // ---- set up the lanes that got work ----
    if (fresh) {
      const int sb = slots + tslot * (PS_WORDS * 4);
      const int pend = (int)lds_ld(c.sb, sb + PS_PEND * 4);
      const int budget = (int)lds_ld(c.sb, sb + PS_BUDGET * 4);
      const int nintro = (int)lds_ld(c.sb, sb + PS_NINTRO * 4);
      const int ch = tchild == ADV_CHILD ? (int)lds_ld(c.sb, sb + PS_CHOSEN * 4) : tchild;
      probe_child(ch, base[pend], budget, nintro, a.pb, tval, tnb, tni);
      for (int wd = 0; wd < E13_TW; ++wd) lds_st(c.sb, c.tbl + wd * 256, lds_ld(c.sb, sb + (PS_TAPE + wd) * 4));
      build_rows(c, sbytes + sb + PS_TAB * 4, nq, pend, tval, 1, nullptr);
      const int r0 = sb + PS_REG * 4;
      L.q = (int)lds_ld(c.sb, r0);
      L.row = c.rbl + (int)lds_ld(c.sb, r0 + 4) * 256;
      L.t = (int)lds_ld(c.sb, r0 + 8);
      L.nt = (int)lds_ld(c.sb, r0 + 12);
      L.stage = (int)lds_ld(c.sb, r0 + 16);
      L.st0 = (int)lds_ld(c.sb, r0 + 20);
      L.q0 = (int)lds_ld(c.sb, r0 + 24);
      L.flags = (int)lds_ld(c.sb, r0 + 28);
      L.ones = (int)lds_ld(c.sb, r0 + 32);
      L.hb = GATE == GATE_HASH ? lds_ld64(c, r0 + 36) : 0ull;
      L.ch0 = lds_ld64(c, r0 + 44);
      L.ch1 = lds_ld64(c, r0 + 52);
      L.ch2 = lds_ld64(c, r0 + 60);
      L.flags |= F_DM; // not known: the first step tests the gate
      arm_stage(L, c);
      t0 = L.t;
      active = 1;
    }
    const uint64_t act = ballot(active);
    if (act == 0) {
      if (exhausted && busy == 0) break;
      if (!progress) { // cannot happen: a slot that holds a probe always has a next step
        if (lane == 0) atomicOr(&a.ctr->overflow, 8u);
        break;
      }
      continue;
    }
    w_rounds++;
    w_slots += (uint64_t)__popcll(act);
    // ---- a round of steps ----
    int ev = EV_NONE, evcell = -1;
    const bool solo = GATE == GATE_COUNT && a.solo > 0 && __popcll(act) == 1;
    if (solo) { // the whole wave runs the lone lane's loop (its block memo takes idle lanes' tapes)
      solo_memo_reset(active, c, act, fills, solo_last);
      ev = run_solo(L, c, a.solo, evcell, __ffsll((long long)act) - 1);
      if (!active) ev = EV_NONE;
    } else if (active) {
      ev = run_slice<GATE, false, false>(L, c, a.slice, 0, evcell);
    }
    // ---- finished runs ----
    if (active && ev != EV_NONE) {
      active = 0;
      const int sb = slots + tslot * (PS_WORDS * 4);
      const uint32_t dt = (uint32_t)(L.t - t0);
      if (tchild == ADV_CHILD) { // the drawn child again: the next node (e13search.c:305-311)
        adv_steps += dt;
        if (ev == EV_STOP) {
          if (probe_next_node(c, L, sbytes, sb, evcell, tval, tnb, tni, base)) {
            tchild = 0;
            t0 = L.t;
            active = 1;
          }
        } else { // cannot happen: the drawn child stopped at this cell the first time
          atomicOr(&a.ctr->overflow, 16u);
          lds_st(c.sb, sb + PS_PH * 4, PH_EMPTY);
        }
      } else if (lds_ld64(c, sb + PS_EXIST * 4) == (1ull << tchild)) {
        // the level's only child: its sums, the draw among one live child and
        // the next node are this lane's (e13search.c:289-311)
        my_steps += dt;
        const bool ended = ev == EV_DEAD || ev == EV_SURV;
        probe_single_add(c, sb, (int)dt, ended, ev == EV_SURV);
        if (ended) { // e13search.c:303
          const uint32_t p = lds_ld(c.sb, sb + PS_P * 4);
          for (int j = 0; j < 4; ++j)
            a.out[4 * (size_t)p + j] = __builtin_bit_cast(double, lds_ld64(c, sb + (PS_SUM + 2 * j) * 4));
          lds_st(c.sb, sb + PS_PH * 4, PH_EMPTY);
        } else {
          uint64_t rng = lds_ld64(c, sb + PS_RNG * 4);
          (void)probe_mix(rng); // splitmix(rng) % 1; w *= 1
          lds_st64(c, sb + PS_RNG * 4, rng);
          if (probe_next_node(c, L, sbytes, sb, evcell, tval, tnb, tni, base)) {
            tchild = 0;
            t0 = L.t;
            active = 1;
          }
        }
      } else {
        my_steps += dt;
        lds_st(c.sb, sb + (PS_DT + tchild) * 4, dt);
        const uint32_t bit = 1u << (tchild & 31);
        const int wo = tchild >> 5;
        if (ev == EV_DEAD || ev == EV_SURV) atomicOr(&smem[(sb >> 2) + PS_ENDED + wo], bit);
        if (ev == EV_SURV) atomicOr(&smem[(sb >> 2) + PS_SURV + wo], bit);
        atomicSub(&smem[(sb >> 2) + PS_LEFT], 1u);
      }
    }
    wave_lds_sync();
  }
  const uint64_t s_steps = wave_sum64(my_steps), s_adv = wave_sum64(adv_steps);
  if (lane == 0) {
    atomicAdd((unsigned long long *)&a.ctr->steps, (unsigned long long)s_steps);
    atomicAdd((unsigned long long *)&a.ctr->cyc[0], (unsigned long long)s_adv); // re-run steps
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
  return (size_t)nw * (ROWS_BYTES + TAPE_BYTES + 64) + (gate == GATE_HASH ? KEYB_WORDS * 8 : 0) + CKEY_BYTES;
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

int e13k_launch_seed_ids(E13Node *seeds, uint32_t n, hipStream_t s) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_seed_ids, dim3((n + 255) / 256), dim3(256), 0, s, seeds, n);
  return (int)hipGetLastError();
}

int e13k_launch_chain_pop(E13Chain *ch, E13Counters *ctr, hipStream_t s) {
  hipLaunchKernelGGL(k_chain_pop, dim3(1), dim3(64), 0, s, ch, ctr);
  return (int)hipGetLastError();
}

int e13k_launch_chain_push(E13Chain *ch, E13Node *stack, uint64_t stack_cap, const E13Node *out, uint32_t out_cap,
                           const E13Counters *ctr, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_chain_push, dim3(grid), dim3(256), 0, s, ch, stack, stack_cap, out, out_cap, ctr);
  return (int)hipGetLastError();
}

// tw: tape words per lane in LDS (E13_TW: the whole window; fewer: the
// narrow form, count gate only).
size_t e13k_search_lds_bytes(int wg_threads, int gate, int np, int tw) {
  const int nw = wg_threads / WAVE;
  return (size_t)nw * (ROWS_BYTES + (size_t)tw * 256 + STG_WORDS * 4) + (gate == GATE_HASH ? KEYB_WORDS * 8 : 0) +
         (size_t)np * E13_PCNT * 8 + (size_t)np * PT_WORDS * 4 + PT_DL_WORDS * 4 + CKEY_BYTES +
         (size_t)nw * (tw < E13_TW ? 0 : CODES_BYTES); // the short-tape form decodes the rows instead
}

// Workgroups of k_search that fit on one CU at once (registers, LDS, wave
// slots); -1 when the runtime cannot tell.
int e13k_search_max_wgs(int wg_threads, int gate, int np, int tw) {
  const size_t lds = e13k_search_lds_bytes(wg_threads, gate, np, tw);
  int n = 0;
  hipError_t e;
  if (gate == GATE_HASH) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_search<GATE_HASH, false>, wg_threads, lds);
  else if (tw < E13_TW) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_search<GATE_COUNT, true>, wg_threads, lds);
  else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_search<GATE_COUNT, false>, wg_threads, lds);
  return e == hipSuccess ? n : -1;
}

int e13k_launch_search(const E13SearchArgs *a, int grid, int wg_threads, int gate, hipStream_t s) {
  const int tw = gate == GATE_HASH ? E13_TW : a->tw;
  if (tw < E13_TW_MIN || tw > E13_TW) return (int)hipErrorInvalidValue;
  const size_t lds = e13k_search_lds_bytes(wg_threads, gate, a->np, tw);
  if (gate == GATE_HASH) hipLaunchKernelGGL((k_search<GATE_HASH, false>), dim3(grid), dim3(wg_threads), lds, s, *a);
  else if (tw < E13_TW && (a->phase_cycles || a->trace != nullptr))
    hipLaunchKernelGGL(k_search_clocked, dim3(grid), dim3(wg_threads), lds, s, *a);
  else if (tw < E13_TW) hipLaunchKernelGGL((k_search<GATE_COUNT, true>), dim3(grid), dim3(wg_threads), lds, s, *a);
  else hipLaunchKernelGGL((k_search<GATE_COUNT, false>), dim3(grid), dim3(wg_threads), lds, s, *a);
  return (int)hipGetLastError();
}

int e13k_launch_batch(const E13BatchArgs *a, int grid, int wg_threads, int gate, int refill, hipStream_t s) {
  const size_t lds = e13k_lds_bytes(wg_threads, gate);
  if (gate == GATE_HASH) {
    if (refill) hipLaunchKernelGGL((k_batch<GATE_HASH, 1>), dim3(grid), dim3(wg_threads), lds, s, *a);
    else hipLaunchKernelGGL((k_batch<GATE_HASH, 0>), dim3(grid), dim3(wg_threads), lds, s, *a);
  } else {
    if (refill) hipLaunchKernelGGL((k_batch<GATE_COUNT, 1>), dim3(grid), dim3(wg_threads), lds, s, *a);
    else hipLaunchKernelGGL((k_batch<GATE_COUNT, 0>), dim3(grid), dim3(wg_threads), lds, s, *a);
  }
  return (int)hipGetLastError();
}

// The probe kernel: the expansion's LDS layout (rows, tapes, block keys, 64
// bytes per wave) and each wave's PK probe slots.
size_t e13k_probe_lds_bytes(int wg_threads, int gate) {
  return e13k_lds_bytes(wg_threads, gate) + (size_t)(wg_threads / WAVE) * PK * PS_WORDS * 4;
}

int e13k_probe_max_wgs(int wg_threads, int gate) {
  const size_t lds = e13k_probe_lds_bytes(wg_threads, gate);
  int n = 0;
  const hipError_t e = gate == GATE_HASH
                           ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_probe<GATE_HASH>, wg_threads, lds)
                           : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_probe<GATE_COUNT>, wg_threads, lds);
  return e == hipSuccess ? n : -1;
}

int e13k_launch_probe(const E13ProbeArgs *a, int grid, int wg_threads, int gate, hipStream_t s) {
  const size_t lds = e13k_probe_lds_bytes(wg_threads, gate);
  if (gate == GATE_HASH) hipLaunchKernelGGL((k_probe<GATE_HASH>), dim3(grid), dim3(wg_threads), lds, s, *a);
  else hipLaunchKernelGGL((k_probe<GATE_COUNT>), dim3(grid), dim3(wg_threads), lds, s, *a);
  return (int)hipGetLastError();
}

} // extern "C"

// ---------------------------------------------------------------------------
// Device scratch of the estimator entry point that allocates nothing
// (e13_estimate_static): per-probe sums (E13_SCRATCH_PROBES x 4 doubles), then
// the launch counters and the problem record.  Each GPU holds its own copy at
// its own address, found once per device with hipGetSymbolAddress; host code
// uses only that address (the array's name is not a device address there).
// Every call on a GPU shares it: calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[E13_SCRATCH_BYTES];

namespace {
#define E13_SCRATCH_DEVICES 64
void *g_scratch_ptr[E13_SCRATCH_DEVICES];
int g_scratch_err[E13_SCRATCH_DEVICES]; // 0 not looked up yet, 1 found, -1 the lookup failed
}

__global__ void __launch_bounds__(256) k_zero_words(uint32_t *p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

extern "C" {

// This device's scratch address (the device must be current); 0 or a HIP error.
int e13k_scratch_addr(int device, void **ptr) {
  if (device < 0 || device >= E13_SCRATCH_DEVICES) return (int)hipErrorInvalidDevice;
  if (g_scratch_err[device] == 0) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
      g_scratch_err[device] = -1;
      return (int)e;
    }
    g_scratch_ptr[device] = p;
    g_scratch_err[device] = 1;
  }
  if (g_scratch_err[device] < 0) return (int)hipErrorInvalidSymbol;
  *ptr = g_scratch_ptr[device];
  return 0;
}

// Zero nwords words at p with a fill kernel (no memset node, safe under capture).
int e13k_launch_zero(void *p, uint32_t nwords, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_words, dim3(8), dim3(256), 0, s, (uint32_t *)p, nwords);
  return (int)hipGetLastError();
}

} // extern "C"
