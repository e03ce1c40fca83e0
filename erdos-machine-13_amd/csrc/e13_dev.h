/* e13_dev.h -- records shared by the C host and the gfx950 kernels.
 *
 * Tape layout.  Bit q of a lane's tape holds cell q - 63, so the live window
 * of the reference (head in -63..1298, e13search.c:173) is q in [0, 1361]
 * and one unsigned compare checks it.  Cells -63..-3 (the counter, hashed
 * into hc at e13search.c:170) are q 0..60; cells >= -2 (the block, hashed
 * into hb) are q >= 61.  43 words of 32 bits hold the window.
 *
 * Table codes (one byte per cell, cell = 2*state + symbol):
 *   0..59  a move, (next << 2) | (dir << 1) | write, dir 1 = R (e13search.c:72)
 *   61     an unread cell that may still be edited (delete mode, budget > 0)
 *   62     halt
 *   63     a free cell not yet chosen (FORCED, e13search.c:66)
 */
#ifndef E13_DEV_H
#define E13_DEV_H

#include <stdint.h>

#define E13_NS 9
#define E13_TW 43           /* tape words per machine */
#define E13_TW_MIN 18       /* shortest lane tape of the narrow form: the stage-9 block (q <= 573) fits */
#define E13_TW_PAD 44       /* words stored in a node record */
#define E13_QMAX 1361       /* largest live q */
#define E13_QOFF 63         /* q = cell + 63 */
#define E13_QBLK 61         /* first q of the block region (cell -2) */
#define E13_MAXCELLS 32
#define E13_MAXNQ 15        /* states: codes must stay below 60 */

#define E13_C_EDIT 61
#define E13_C_HALT 62
#define E13_C_FREE 63

/* A stopped machine: the state right before the read of cell `pend`, or
   (pend < 0) a machine paused by the per-launch step cap.  304 bytes = 76
   words; the word offsets are used by the kernels' cooperative copies. */
typedef struct E13Node {
  int32_t q, st, t, nt;            /* head (q), state, steps from the blank tape, next target stage */
  int32_t stage, st0, q0, flags;   /* last accepted stage, stage-1 state and head; flags: bit0 abs_ok,
                                      bit1 rel_ok, bits 2..4 cs[0..2] */
  int32_t ones, budget, nintro, pend; /* ones at q >= 61, edits left, symmetric states in use, cell to bind */
  int32_t prob, depth, nbound, next_slot; /* problem index, binding depth, cells bound so far, and the
                                             child slots not yet handed out: first | end << 16 (end 0 =
                                             all slots; a split node carries part of the range) */
  uint64_t hb;                     /* block hash (hash-gate build only) */
  uint64_t ch[3];                  /* counter hashes at the accepted stages */
  uint8_t tab[E13_MAXCELLS];       /* word 24 */
  uint32_t tape[E13_TW];           /* word 32 */
  uint32_t seed;                   /* word 75: the frontier node this node descends from (cost logging) */
} E13Node;
#define E13_NODE_WORDS 76
#define E13_NODE_W_SLOTS 15 /* word of next_slot */
#define E13_NODE_W_TAB 24
#define E13_NODE_W_TAPE 32
#define E13_NODE_W_SEED 75

/* Survivor as the kernel records it. */
typedef struct E13SurvRec {
  uint8_t tab[E13_MAXCELLS];
  int32_t st0, head0, abs_ok, rel_ok, steps, prob;
  uint64_t ch[3];
} E13SurvRec;

/* Result of one complete table run from the blank tape. */
typedef struct E13LeafRec {
  int32_t result; /* -1 dead, -2 survivor (e13search.c:64-65) */
  int32_t stage, steps, st0, head0, abs_ok, rel_ok, pad;
  uint64_t ch[3];
} E13LeafRec;

/* Device counters.  Totals are 64-bit; the per-launch work cursor and slot
   counters are 32-bit. */
typedef struct E13Counters {
  uint64_t leaves, steps, survivors;
  uint64_t hist[E13_NS + 1];
  uint64_t lane_slots;  /* lanes x rounds executed (for lane utilisation) */
  uint64_t rounds;      /* wave rounds */
  uint32_t item;        /* next work item of the current launch */
  uint32_t nout;        /* children written by the current launch */
  uint32_t nsurv;       /* survivor slots taken (may exceed the capacity) */
  uint32_t overflow;    /* set when a child did not fit the output buffer */
  uint64_t cyc[4];      /* wave cycles (s_memtime) in refill, stepping, finishing, scheduling */
} E13Counters;

/* Per-problem constants of the expansion kernel. */
typedef struct E13Problem {
  int32_t nq, nslot, X, sym_lo, sym_hi;
  int32_t dl[E13_NS + 1];   /* per-stage deadlines, index 1..9 */
  int8_t base[E13_MAXCELLS]; /* codes after setup(): move, -1 halt, -2 free */
} E13Problem;

/* Device-resident cursor of the host-loop forms' node stack, for launch chains
   captured in a hipGraph: each iteration pops `nin` nodes from `top`, expands
   them, and pushes the children back, with no host round trip. */
typedef struct E13Chain {
  uint64_t top;      /* nodes on the stack */
  uint64_t base;     /* first node of the current batch */
  uint64_t peak;     /* largest top seen */
  uint32_t nin;      /* nodes in the current batch */
  uint32_t batch;    /* most nodes per batch */
  uint32_t error;    /* 1: the stack or the output buffer overflowed */
  uint32_t iters;    /* iterations that found work */
  uint64_t pad[2];
} E13Chain;

typedef struct E13ExpandArgs {
  const E13Node *in;      /* ignored when chain is set: in = stack + chain->base */
  E13Node *out;
  E13Counters *ctr;
  E13SurvRec *surv;
  const uint64_t *key;    /* Zobrist keys KEY[0..1363] */
  const uint64_t *tgt;    /* TGT[0..9] */
  const E13Problem *pb;   /* device copy of the problem constants */
  uint32_t nin, out_cap, surv_cap, nslot;
  int32_t slice, cap, refill_tries, nq;
  uint32_t chunk, pad0, pad1, pad2; /* work items a wave claims per atomic */
  E13Chain *chain;        /* set: nin and in come from the device cursor */
  const E13Node *stack;   /* the stack the chain cursor indexes */
} E13ExpandArgs;

/* Scheduling state of the persistent search (one per search). */
typedef struct E13Sched {
  uint64_t pad1[2];
  int32_t busy;         /* waves holding work, plus deliveries not yet taken */
  int32_t error;        /* nonzero: abort (1 private stack full, 2 incomplete delivery, 3 idle watchdog,
                           4 LDS base, 5 not every workgroup resident: no work was done) */
  int32_t donations, pool_takes; /* nodes delivered / taken */
  int32_t exited;       /* waves that have left */
  int32_t halt;         /* narrow form: nonzero once the list of wide nodes is full: every wave stops
                           cleanly, as at a stop request */
  uint32_t wide_n;      /* narrow form: nodes put on (or refused by) the list of wide nodes */
  uint32_t wide_cap;    /* narrow form: capacity of that list */
  E13Node *wide;        /* narrow form: nodes whose head or written cells lie past the lane's LDS tape,
                           with their slot cursors, for a launch with the whole window */
  int32_t started;      /* waves that have started (the check that every workgroup is resident) */
  int32_t pad3;
} E13Sched;

/* A wave's mailbox.  A donor that won the wave's bit in the hungry map takes
   the mailbox's next delivery number from `seq` (an atomic add), writes the
   node as 64-bit elements (delivery tag << 32 | node word), then sets
   flag = tag, with tag = seq mod (2^32 - 1) + 1.  The wave polls only its own
   flag; a flag that differs from the last tag it took is a delivery, and the
   node is complete when every element carries that tag (stale words from an
   earlier delivery to this mailbox carry an earlier number).  Numbers are per
   mailbox, so two deliveries to one wave carry the same tag only 2^32 - 1
   deliveries apart.  flag = E13_MAIL_DONE: the search is over.  Tags are
   never 0. */
#define E13_MAIL_DONE 0xFFFFFFFFFFFFFFFFull
typedef struct E13Mail {
  uint64_t flag;
  uint64_t seq;   /* deliveries numbered so far */
  uint64_t pad[6];
  uint64_t w[E13_NODE_WORDS];
} E13Mail;

/* Static device scratch of the estimator entry point that allocates nothing:
   per-probe sums for up to E13_SCRATCH_PROBES probes (4 doubles each), then the
   launch counters (E13Counters at E13_SCRATCH_CTR), the problem record
   (E13Problem at E13_SCRATCH_PB) and the root node (E13Node at E13_SCRATCH_ROOT). */
#define E13_SCRATCH_PROBES (1 << 16)
#define E13_SCRATCH_CTR (E13_SCRATCH_PROBES * 32)
#define E13_SCRATCH_PB (E13_SCRATCH_CTR + 1024)
#define E13_SCRATCH_ROOT (E13_SCRATCH_PB + 1024)
#define E13_SCRATCH_BYTES (E13_SCRATCH_ROOT + 512)

#define E13_MAXPROB 16
#define E13_PCNT 16       /* per-problem counters: leaves, steps, survivors, hist[0..9] */

typedef struct E13SearchArgs {
  E13Node *stacks;        /* nwaves x stk_cap private stacks (rings; stk_cap a power of two) */
  E13Mail *mail;          /* one mailbox per wave */
  uint64_t *hungry;       /* hungry map: bit w set = wave w asks for a node (hungry_words words) */
  E13Sched *sch;
  const E13Problem *pbs;  /* np problems */
  uint64_t *pcnt;         /* np x E13_PCNT */
  E13Counters *ctr;       /* survivor slots, rounds, lane slots */
  E13SurvRec *surv;
  const uint64_t *key;
  const uint64_t *tgt;
  int32_t np, slice, stk_cap, hungry_words;
  int32_t refill_tries, donate_every, pool_min_idle, total_waves;
  uint32_t surv_cap, spin_limit;
  uint32_t idle_limit;    /* idle watchdog, in ticks of the 100 MHz constant clock */
  uint32_t sched_every;   /* rounds between reads of the hungry map while no wave is hungry */
  const E13Node *seeds;   /* nseeds frontier nodes dealt round-robin to the waves' stacks at start */
  uint32_t nseeds;
  int32_t idle_naps, mail_every, trace_cap;
  int32_t solo;           /* > 0: steps per round of the scalar loop for a wave's lone running lane */
  int32_t evict_after;    /* > 0: a machine this many steps from the blank tape may be handed to a hungry wave */
  int32_t cut_live;       /* > 1: a round ends once fewer lanes than this still run while the wave has nodes */
  int32_t phase_cycles;   /* nonzero: count wave clock cycles per phase (ctr->cyc) */
  int32_t skip_settled;   /* nonzero: the step loop tests the stage-1 gate only after a step that changed
                             the ones count (Lane.dm) */
  int32_t tw;             /* tape words per lane in LDS: E13_TW, or fewer (the narrow form; its list of wide
                             nodes is in sch) */
  uint32_t nseed_cnt;     /* cost logging: seeds with counters (0: off) */
  uint64_t *seed_cnt;     /* cost logging: per seed, leaves and steps of its subtree */
  uint64_t *trace;        /* optional: trace_cap words per wave, word 0 = events logged, then
                             (realtime << 16 | aux << 8 | event) */
  const uint32_t *stop;   /* optional (clean stop): a flag in host memory; once it is nonzero every wave
                             pauses its running machines as resumable nodes on its stack and leaves */
  int64_t *wave_state;    /* with stop: per wave, its stack range [bot, top) and the tag of the last
                             delivery it took from its mailbox, written when it leaves */
} E13SearchArgs;

typedef struct E13BatchArgs {
  const uint8_t *tabs;    /* n x 32 codes (0..59 move, 62 halt) */
  const int32_t *caps;    /* n caps replacing every stage deadline, or NULL */
  E13LeafRec *out;
  E13Counters *ctr;
  const uint64_t *key;
  const uint64_t *tgt;
  const int32_t *dl;      /* device copy of the stage deadlines dl[0..9] */
  int32_t n, nq, slice;
  uint32_t chunk;         /* tables a wave claims per atomic (refilling form) */
  int32_t solo;           /* > 0: steps per round of the scalar loop for a wave's lone running lane */
  int32_t pad_b;
} E13BatchArgs;

/* Probe estimator (probe(), e13search.c:256-313): one wave walks one probe. */
typedef struct E13ProbeArgs {
  const E13Problem *pb;   /* the problem (base codes, symmetry) */
  const uint64_t *key;
  const uint64_t *tgt;
  double *out;            /* n x 4: leaves, nodes, steps, survivors of each probe (weighted) */
  E13Counters *ctr;       /* item = next probe of this launch; steps (children), cyc[0] = steps of the
                             re-runs that make the next node of a path, rounds, lane slots */
  uint64_t seed;          /* probe p0 + i draws from the stream of (seed, p0 + i) */
  uint64_t p0;
  uint32_t n;             /* probes in this launch */
  int32_t nq, slice;
  int32_t solo;           /* > 0: steps per round of the scalar loop for a wave's lone running lane */
  int32_t pad_p;
  const E13Node *roots;   /* probe p0 + i starts at roots[(p0 + i) / per_root]: a stopped machine with its
                             pending cell (a root without one, paused at a cap, gets sums of -1) */
  uint64_t per_root;
} E13ProbeArgs;

/* Deep stage check / counter re-check of survivors on the GPU (k_deep): one
   wave per survivor, its whole tape in LDS (W words, cell k at bit k + left),
   resumable over launches of a bounded number of loop iterations. */
#define E13_DEEP_MAXQ 32       /* states of a survivor table */
typedef struct E13DeepState {
  int64_t t;                   /* steps so far */
  int64_t limit;               /* step limit of the stage sought */
  int64_t fail_t;              /* the step count of the failure */
  int32_t q, st, ones, n;      /* head (cell + left), state, ones at cells >= -2, stage sought */
  int32_t minq;                /* leftmost head so far */
  int32_t status;              /* 0 running, 1 finished, 2 left the tape held in LDS (re-check: the host runs it),
                                  3 dynamic LDS not at address 0 */
  int32_t fail;                /* 0 every stage reached, 1 step limit, 2 off the tape, 3 free cell read, 4 halt read */
  int32_t fail_n, fail_q, fail_st, fail_sym, pad;
} E13DeepState;

typedef struct E13DeepArgs {
  const uint32_t *rows;        /* per survivor E13_DEEP_MAXQ words: entry(state, 0) | entry(state, 1) << 16 */
  const int32_t *param;        /* per survivor: st0, head0 + left, relative start (0/1), 0 */
  E13DeepState *state;
  uint32_t *tape;              /* per survivor W words */
  uint32_t *snap;              /* per survivor (nmax + 1) x (left / 32) words: the counter side of the tape at
                                  each stage start */
  const int32_t *list;         /* the survivors of this launch, one per workgroup */
  int32_t nmax, W, left, mode; /* mode 0: deep.c (window -left+2 .. 2^n+2000); 1: recheck.c (the tape held) */
  int64_t kmul;                /* step limit of stage n: (stage n-1's step) + kmul * 4^n + 100000 */
  int64_t budget;              /* loop iterations per launch */
} E13DeepArgs;

#endif
