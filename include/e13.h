/* e13.h -- C ABI of the MI355X (gfx950) TNF-enumeration + simulate-and-filter
 * library, libe13.so.  Plain C types only.
 *
 * The reference keeps this path in static functions of src/e13search.c plus a
 * CuPy kernel (src/e13gpu.py); it has no FFI.  Each entry point below names
 * the reference function or call it replaces.
 *
 * Conventions: functions return 0 or a negative E13_E* code and never exit;
 * the caller owns every host buffer, the context owns device memory; one
 * context per host thread (not re-entrant), one HIP stream inside.
 */
#ifndef E13_H
#define E13_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define E13_NS 9
#define E13_MAXCELLS 32

/* Error codes. */
#define E13_OK 0
#define E13_EINVAL (-1)     /* bad argument or shape */
#define E13_EHIP (-2)       /* HIP runtime error */
#define E13_ENOMEM (-3)     /* device or host allocation failed */
#define E13_EFORCED (-4)    /* forced cells exceed FMAX (e13search.c:353) */
#define E13_ESYM (-5)       /* a symmetric state is named by a fixed cell or has one (e13search.c:346) */
#define E13_ESYMSTART (-6)  /* start state in the symmetric states (e13search.c:349) */
#define E13_ESTACK (-7)     /* node stack capacity exceeded (e13gpu.py:420-421) */
#define E13_ENODEV (-8)     /* no HIP device */
#define E13_ESTOPPED (-9)   /* a clean stop: the checkpoint is saved (e13gpu.py:398-409, exit code 3) */
#define E13_ECKPT (-10)     /* checkpoint unreadable, damaged or written for another search (e13gpu.py:360-361) */
/* Aborts of the persistent search's scheduler (no reference counterpart: the
   reference's host loop has no device-side scheduling). */
#define E13_ELOST (-11)     /* a mailbox delivery never arrived complete */
#define E13_EWATCHDOG (-12) /* a wave waited longer than the idle watchdog for work or the end mark */
#define E13_ELDS (-13)      /* dynamic LDS does not start at LDS address 0 */
#define E13_ESCRATCH (-14)  /* e13_estimate_static: more probes than its static scratch holds */
#define E13_ERESIDENT (-15) /* not every workgroup of the persistent search became resident at once */

/* Result codes of a run (e13search.c:64-65). */
#define E13_DEAD (-1)
#define E13_SURV (-2)

/* Table codes: (next << 2) | (dir << 1) | write, dir 1 = R; -1 halt; -2 a
   cell still to be chosen (FORCED) -- e13search.c:66, 72. */
#define E13_HALT (-1)
#define E13_FORCED (-2)

/* A search after setup()/set_sym() (e13search.c:315-355, 422-426). */
typedef struct {
  int32_t nq;       /* states after the deletion */
  int32_t X;        /* deleted base state, -1 none (printed as X=) */
  int32_t budget;   /* edits left after the forced cells (0 in free mode) */
  int32_t sym_lo, sym_hi, sym_n0; /* symmetric states, renumbered; -1 = off */
  int32_t nforced;  /* rewritten rules into X */
  int8_t tab[E13_MAXCELLS];
} e13_problem;

typedef struct {
  uint64_t leaves, steps, survivors;
  uint64_t dead_at_stage[E13_NS + 1];
} e13_counts;

typedef struct {
  int8_t tab[E13_MAXCELLS]; /* -1: halt, or a free cell never chosen */
  uint32_t read_mask;       /* bit i: cell i read before the stage-9 start ('?' when 0) */
  int32_t st0, head0, abs_ok, rel_ok;
  int64_t steps;
  uint64_t ch[3];           /* counter hashes at the stage starts, classes 0..2 */
  int32_t recheck_ok;       /* 1: the plain unhashed re-check agrees */
  int32_t prob;             /* index of its problem (e13_search_many) */
} e13_survivor;

typedef struct {
  int32_t result; /* E13_DEAD or E13_SURV */
  int32_t stage, steps, st0, head0, abs_ok, rel_ok, event;
  uint64_t ch[3];
} e13_leaf;

/* Launch and schedule settings.  Any field <= 0 takes the default (or the
   autotuned value when one is stored for this device). */
typedef struct {
  int32_t form;        /* 1 plain: one machine per lane, one launch per batch of nodes, host loop;
                          2 batched: the same host loop with persistent lanes refilled in the launch;
                          3 persistent: one launch for the whole search, device work queue */
  int32_t slice;       /* steps per round between lane refills */
  int32_t wg_threads;  /* 64 .. 256 */
  int32_t wgs_per_cu;  /* persistent grid = CUs x this */
  int32_t cap;         /* step cap per machine per launch */
  int32_t batch;       /* input nodes per launch */
  int32_t gate;        /* 1 ones-count gate, 2 reference Zobrist hash gate */
  int32_t refill_tries;
  int64_t stack_nodes; /* node stack capacity */
  /* persistent form (3) scheduling */
  int32_t seed_per_wave;  /* breadth-first frontier per wave before the launch; 0 = start from the roots */
  int32_t sched_every;    /* rounds between reads of the hungry map while no wave is hungry */
  int32_t donate_max;     /* nodes a wave may deliver per read of the hungry map */
  int32_t idle_naps;      /* an idle wave's pause between mailbox polls, in units of s_sleep(8) */
  int32_t mail_every;     /* rounds between mailbox polls of a wave whose lanes still run */
  int32_t trace_events;   /* > 0: log up to this many scheduling events per wave of a search call's
                             first launch (e13_get_trace) */
  /* host-loop forms (1, 2) */
  int32_t graph_chain;    /* > 0: iterations (pop, expand, push) per captured hipGraph; 0 = one launch
                             and one host round trip per iteration */
  /* persistent form: the frontier and the cost model (e13_set_cost_model) */
  int32_t cost_log;       /* 1: log each frontier node's features and subtree leaves/steps (e13_get_cost_log) */
  int32_t learned;        /* 1: split and order the frontier by the installed cost model; 2: a seeded
                             random order (a check); 3: split and order it by probe estimates of each
                             node's subtree (the reference's probe() from the node, 16 probes: the
                             non-learned baseline); 0 or negative: the static split */
  /* tape placement: a wave whose only running lane has nothing to refill keeps that lane's
     rule rows and current tape word in registers and steps it on the scalar unit */
  int32_t solo_steps;     /* > 0: steps per such round (registers); negative: off (the LDS loop only) */
  int32_t evict_after;    /* > 0: while waves are hungry, a wave with nothing to give hands over machines
                             this many steps from the blank tape, and a wave whose only running machine is
                             one keeps it alone (register loop); negative: off */
  int32_t probe_slice;    /* probe estimator (e13_estimate): steps per round between hand-outs of child runs */
  int32_t cut_live;       /* persistent form: > 1: a round of the multi-lane loop ends once fewer than this many
                             of the wave's lanes still run while the wave has nodes to hand out, so the idle lanes
                             are refilled sooner (the running ones go on next round); negative: off */
  int32_t phase_cycles;   /* persistent form: > 0: count wave clock cycles per phase into e13_stats.cycles (a
                             diagnostic that costs about 2%); negative: off */
  int32_t nt1_skip;       /* persistent form, step loop: 1: test the stage-1 gate only after a step that changed
                             the ones count (before stage 1 the block can become exact in no other way), so a
                             machine whose two ones sit in the wrong cells does not leave the loop at every
                             step; 2: test it at every step; 0 or negative: 1 when some problem has an edit
                             budget (delete mode), else 2.  Never changes results. */
  int32_t tape_words;     /* persistent form: tape words per lane in LDS.  0 or negative (automatic): when the
                             whole tape (43 words, the window of cells -63..1298) does not let wgs_per_cu
                             workgroups fit a CU, the widest shorter tape with which the most of them do; 18..42:
                             that tape; 43: the whole tape.  With a shorter tape, machines whose head goes past
                             it are paused and finish in a second launch with the whole tape.  Never changes
                             results. */
  int32_t wide_nodes;     /* capacity of the list of those machines (nodes; 0: 2^19); when it fills, the
                             launch stops cleanly and the rest of the search runs on the whole tape */
} e13_tuning;

typedef struct {
  double kernel_ms;      /* summed HIP-event time of the expansion/batch kernels */
  int64_t launches;
  int64_t rounds;        /* wave rounds */
  int64_t lane_slots;    /* active lanes summed over rounds */
  int64_t peak_stack;
  int32_t slice, wg_threads, grid, gate, form;
  int32_t profile;       /* settings profile of the last persistent search (E13_PROFILE_*) */
  int64_t donations, pool_takes; /* persistent form: nodes moved between waves */
  int64_t cycles[4];             /* wave clock cycles spent refilling lanes, stepping, finishing
                                    machines, and on the work pool / idle (summed over waves; the
                                    persistent form counts them with tuning.phase_cycles > 0 only) */
  int64_t graphs;                /* hipGraph launches (graph_chain > 0) */
  double search_ms;              /* the persistent kernel alone (k_search): HIP-event time, */
  int64_t search_launches;       /* launches, */
  int64_t search_steps;          /* and the steps simulated inside it */
  int64_t frontier;              /* nodes the last search's parallel phase started from: the breadth-first
                                    frontier of the persistent form (the reference's front=, e13search.c:373),
                                    1 (the root) for the host-loop forms */
  int32_t tape_words;            /* persistent form: lane tape words of the last search's first launch */
  int32_t pad_s;
  int64_t wide_nodes;            /* machines that went past the short tape (tape_words < 43) */
  int64_t wide_launches;         /* launches with the whole tape after a short-tape launch, */
  double wide_ms;                /* their HIP-event time */
  int64_t wide_steps;            /* and the steps simulated in them (part of search_ms / search_steps) */
} e13_stats;

typedef struct e13_ctx e13_ctx;

/* Device context (no reference counterpart: the reference's CuPy module setup,
   e13gpu.py:318-347). */
int e13_open(int device, const e13_tuning *t, e13_ctx **out);
void e13_close(e13_ctx *ctx);
const char *e13_strerror(int code);
/* Install settings (fields <= 0 keep the current value); they then hold for
   every search (no automatic profile choice).  t = NULL: back to the settings
   e13_open installed, with the automatic choice of the small-search profile. */
int e13_set_tuning(e13_ctx *ctx, const e13_tuning *t);
int e13_get_tuning(e13_ctx *ctx, e13_tuning *t);
int e13_get_stats(e13_ctx *ctx, e13_stats *s);

/* The wave event log of the last persistent search run with trace_events > 0:
   per wave `per_wave` words, word 0 = events logged, then
   (t << 16 | aux << 8 | event) with t in ticks of the GPU's constant 100 MHz
   clock and event 1 seeded, 2 took a pool node while idle, 3 counted busy
   again, 4 idle, 5 exit, 6 gave a node, 7 every 64th round (aux = running lanes,
   +128 when more than two nodes wait on the wave's stack).
   No reference counterpart (a diagnostic of this implementation's scheduler). */
int e13_get_trace(e13_ctx *ctx, uint64_t *buf, uint64_t cap_words, int32_t *waves, int32_t *per_wave);

/* Zobrist tables: KEY[0..1363], TGT[0..9] (e13search.c:437-444). */
int e13_keys(uint64_t *key, uint64_t *tgt);

/* Machine text -> codes; returns the number of states or a negative code
   (load/parse_cell, e13search.c:105-121).  `text` holds the file contents. */
int e13_parse_machine(const char *text, int32_t *base, int max_states);

/* setup() + set_sym() (e13search.c:317-355, 423-426).  freebase == NULL:
   delete mode with budget fmax; else free mode.  sym_lo < 0: no symmetry. */
int e13_setup(const int32_t *base, int nq_base, int X, int fmax, const uint8_t *freebase, int sym_lo, int sym_hi,
              e13_problem *out);

/* The search: drop-in for run_delete() (e13search.c:357-420) and for the GPU
   driver loop of e13gpu.py:381-437.  Counts are exact; survivors beyond
   surv_cap are counted in *n_surv but not stored. */
int e13_search(e13_ctx *ctx, const e13_problem *p, e13_counts *out, e13_survivor *surv, uint64_t surv_cap,
               uint64_t *n_surv);

/* Several problems (e.g. every deleted state X) in one persistent launch: the
   drop-in for a run_delete() loop over X (e13search.c:461-463).  out[k] are
   the counts of ps[k]; survivors carry their problem index. */
int e13_search_many(e13_ctx *ctx, const e13_problem *ps, int np, e13_counts *out, e13_survivor *surv,
                    uint64_t surv_cap, uint64_t *n_surv);

/* Complete tables from the blank tape: run_full() (e13search.c:428-433) over a
   batch.  tabs: n x ncells codes (>= 0 move, -1 halt); caps: NULL for the
   reference deadlines, else one cap per table replacing every stage deadline. */
int e13_simulate_batch(e13_ctx *ctx, const int8_t *tabs, int ncells, int n, const int32_t *caps, e13_leaf *out);

/* Same, on device buffers already resident in HBM (tabs_dev: n x 32 bytes,
   codes 0..59 or 62 halt; caps_dev or NULL; out_dev n records).  Stream-ordered
   on the context stream; returns after the launch is queued. */
int e13_simulate_batch_device(e13_ctx *ctx, const void *tabs_dev, int nq, int n, const void *caps_dev, void *out_dev);
int e13_sync(e13_ctx *ctx);

/* Checkpoints, clean stop and resume: replace the GPU driver's E13_CKPT /
   E13_CKPT_SEC / <file>.stop handling (e13gpu.py:19-23, 349-409).  With a
   path set, e13_search / e13_search_many save the remaining search to
   <path>.a / <path>.b in turn, the pointer file <path> naming the last complete
   one: every every_sec seconds (<= 0: 1800), when the file <path>.stop appears
   (checked at most ten times a second; the file is removed), or after max_launches
   launches of this call (> 0; 0 = no limit).  Host-loop forms (1, 2): between
   launches, the node stack, counters and survivor records.  Persistent form
   (3): the launch is stopped cleanly (every wave pauses its running machines as
   resumable nodes on its stack and leaves), and the waves' stacks plus any
   delivery still in a mailbox are saved with the counters and survivor
   records; on a timed checkpoint the search relaunches from them at once.  A
   stop returns E13_ESTOPPED after the save.  A search started with a
   checkpoint present resumes from it (E13_ECKPT if it belongs to other
   problems, to the other form, or is damaged); results are those of an
   uninterrupted run.  A stop file left from an earlier clean stop is removed
   at the start; the files stay after the search ends.  path NULL or "": off. */
int e13_set_checkpoint(e13_ctx *ctx, const char *path, double every_sec, int64_t max_launches);

/* Probe estimator: replaces the estimate mode of the reference
   (e13search.c:473-510) and its probe() (:256-313), Knuth's estimator with
   every child of a path node run: each probe walks from the root through
   uniformly drawn live children, adding w per child (w = product of the live
   counts above it) to leaves (ended children), nodes, steps (w x the child's
   steps) and survivors.  Probe i (0 <= i < probes) draws from its own
   splitmix64 stream, state splitmix(0x5EED0000 + (seed << 40) + i), so its
   sums depend on (problem, seed, i) alone; the reference shares one stream per
   OpenMP thread (e13search.c:487-491), which ties its draws to the schedule.
   The per-probe sums go to per_probe[4*i .. 4*i+3] when it is given; the totals
   are added in probe order, and mean/se follow e13search.c:503-506. */
typedef struct {
  int64_t probes;
  double sum[4], sq[4];   /* leaves, nodes, steps, survivors: sums of the per-probe values and squares */
  double mean[4], se[4];  /* the estimates and their standard errors */
  uint64_t sim_steps;     /* machine steps of the probes' children, simulated on the device (unweighted) */
  uint64_t rerun_steps;   /* steps of the re-runs of drawn children that make each path's next node */
  double kernel_ms;       /* HIP-event time of the probe kernels */
  int64_t launches, rounds, lane_slots;
} e13_estimate_sum;
int e13_estimate(e13_ctx *ctx, const e13_problem *p, uint64_t seed, int64_t probes, double *per_probe,
                 e13_estimate_sum *out);

/* e13_estimate with no allocation in the call (no hipMalloc/hipFree, no host
   malloc): for callers that cannot allocate -- code under HIP graph capture,
   tight serving loops, harnesses that forbid it.  No reference counterpart.
   The device buffers (per-probe sums, counters, problem record) live in a
   static array of E13_SCRATCH_BYTES on each GPU, found by e13_open; the caller
   passes per_probe (4 * probes doubles, required).  probes <= E13_STATIC_PROBES
   (E13_ESCRATCH otherwise).  Results equal e13_estimate's for the same
   (problem, seed, probes).  Every call on one GPU shares its array: calls that
   overlap on one GPU (two host threads, two contexts) race. */
#define E13_STATIC_PROBES (1 << 16)
int e13_estimate_static(e13_ctx *ctx, const e13_problem *p, uint64_t seed, int64_t probes, double *per_probe,
                        e13_estimate_sum *out);

/* Plain, unhashed re-check of a survivor over stages 1..9 with the search's
   deadlines and window (after recheck.c:57-92): 1 consistent, 0 not.  The
   cells read before the stage-9 start go to *read_mask. */
int e13_recheck(const int8_t *tab, int nq, int st0, int head0, int abs_ok, int rel_ok, int64_t steps,
                uint32_t *read_mask);

/* Deep stage check (replaces src/deep.c): each SURVIVOR line's table run from
   the blank tape to stage nmax (<= 24) on a tape of 4096 + 2^nmax + 4096
   cells, finding the stage starts (state st0, head at head0 or, for rel=1
   abs=0, at head0 - 2 + 2^n, the block -1 .. 2^n - 2 exactly the ones at
   cells >= -2) and printing the counter's ones left of cell -2 at each, as
   deep.c does; it stops at a stage not reached within 64*4^n + 100000 more
   steps, a halt, a never-read free cell ('?---') or off the tape.
   big = 1: src/deep_big.c instead, 4M cells each side, and at the step limit
   the head, state and extent of the ones are reported as it does.
   e13_deep_check_line: one line as survivor `id`, the result line (with its
   newline) into out; returns its length.  e13_deep_check: every line of text
   with table= and st0=, numbered from 1, on `threads` host threads; the result
   lines into out in input order; returns the number of survivors. */
int e13_deep_check_line(const char *line, int id, int nmax, int big, char *out, int outlen);
int e13_deep_check(const char *text, int nmax, int big, int threads, char *out, int64_t outlen);

/* Classification of deep-check output (replaces src/classify_deep.py): the
   survivors text (lines with table=) and the deep output (lines starting with
   "survivor", in the same order) to classify_deep.py's lines -- per kind (the
   first counter change within a class c(n), counters equal across classes,
   the failure with its times and places removed, or CANDIDATE) and start (abs
   or rel), counts most frequent first, then an example line for candidate
   and free-cell kinds.  Returns the number of survivors. */
int e13_classify_deep(const char *survivors, const char *deep, char *out, int64_t outlen);

/* Counter re-check (replaces src/recheck.c): the same stage starts to nmax
   (<= 19) on a tape of 4e6 cells each side, with the exact set of counter ones
   (cells <= -3) at each start; the verdict line "id: ..." is recheck.c's:
   stage not reached (limit 8*4^n + 100000 more steps), off tape, free cell
   read, halt, "counter changes at n=..." (the set differs from the first
   stage of its class c(n): 0 for n <= 2, 1 for 3..8, 2 from 9), "counter
   equal for different c", or "CONSISTENT to n=... (c2 counter ...)" (for nmax
   < 9 the c2 counter is printed empty).  Returns the line's length. */
int e13_recheck_counter_line(const char *line, int id, int nmax, char *out, int outlen);
/* Every line of text with table=, numbered from 1, on `threads` host threads;
   the verdict lines into out in input order; returns the number of lines. */
int e13_recheck_counter(const char *text, int nmax, int threads, char *out, int64_t outlen);

/* The deep check (src/deep.c:15-84, big = 0) and the counter re-check
   (src/recheck.c:22-104) on the GPU of ctx, nmax <= 18: the same result lines
   as e13_deep_check / e13_recheck_counter, in input order; returns the number
   of survivors.  One wave per survivor with its whole tape in LDS, launches of
   a bounded number of loop iterations until every survivor is done.  A
   re-check whose machine leaves the 4096 + 2^nmax + 4096 cells held on the GPU
   (recheck.c's tape is 4e6 cells each side), and a table with a move to a state
   it does not have, run on the host instead. */
int e13_deep_check_gpu(e13_ctx *ctx, const char *text, int nmax, char *out, int64_t outlen);
int e13_recheck_counter_gpu(e13_ctx *ctx, const char *text, int nmax, char *out, int64_t outlen);

/* Independent check with crafted stage starts (replaces src/check.py): the
   stage starts to 11 from the blank tape with the counter depending only on
   c(n) = #{1 <= k < n : 2^k has no digit 2 in base 3}, then from crafted
   starts (the counter of c, a block of m = 2..40 ones) a halt iff c = 2 and m
   has no digit 2, else the start with block 2m and the counter of c or c + 1;
   with no halt cell the halt goes into each never-read cell in turn.  Writes
   check.py's lines for one survivor (the line itself, then one "   halt@...:
   PASS|FAIL ..." line per placement); returns their length. */
int e13_check_survivor(const char *line, char *out, int outlen);

/* The SURVIVOR line of report() (e13search.c:184-200) and the extra HASH line;
   returns the length written (without the NUL). */
int e13_format_survivor(const e13_problem *p, const e13_survivor *s, char *buf, int buflen);
int e13_format_hash(const e13_problem *p, const e13_survivor *s, char *buf, int buflen);

/* Synthetic stress table k: splitmix64 from seed0 + k; 28 cells v = next() % 56;
   halt cell next() % 28; cap floor(10 * 2e4^u), u = (next() >> 11) * 2^-53. */
void e13_synth_table(uint64_t seed0, uint64_t k, int8_t *tab28, int32_t *cap);

/* Totals and an order-independent digest of a batch of complete-table runs.
   Record hash of table k: h = mix(k), then h = mix(h ^ x) for x = result,
   stage, steps, st0, head0, abs_ok | rel_ok << 1 (each as 32-bit two's
   complement), ch0, ch1, ch2, where mix(x) is the splitmix64 output for state
   x; digest = sum of h mod 2^64. */
typedef struct {
  uint64_t n, digest, steps, survivors;
  uint64_t dead_at_stage[10];
  double kernel_ms; /* HIP-event time of the batch kernel */
} e13_batch_sum;

int e13_batch_digest(const e13_leaf *leaves, int64_t n, uint64_t k0, e13_batch_sum *out);

/* The synthetic batching stress test: tables k0 .. k0+n-1 of e13_synth_table,
   each with its cap as every stage deadline, made on the host, copied to HBM,
   and run in one batch launch (form 1: one table per lane; otherwise lanes
   refill from a device cursor as their tables end); returns the totals and
   digest of e13_batch_digest.  No reference counterpart: the workload is
   defined by this build. */
int e13_synth_run(e13_ctx *ctx, uint64_t seed0, uint64_t k0, int64_t n, e13_batch_sum *out);

/* ---- Autotuner (no reference counterpart; the reference adapts only its GPU
   batch size, e13gpu.py:427-430).  Coordinate descent over resident waves per
   CU, steps per round, refill passes, the lone-lane tape placement (LDS loop or
   register loop), frontier per wave, nodes per delivery pass, hungry-map read
   interval, workgroup size, idle poll interval and the step count after which
   a machine goes to an idle wave to run alone, each trial a persistent search
   of ps[0..np) (the median of reps), starting from the profile those problems use.
   Every trial's counts must equal the first's (E13_EINVAL otherwise).  The
   chosen settings are installed in ctx.
   e13_open loads a device's stored settings from $E13_TUNING_FILE or from
   <library dir>/../tuning/<arch>_<CUs>cu.txt (e13_tuning_file) when present,
   and a second profile for small searches from the same path with _small
   before the extension (e13_tuning_file_small); fields the caller sets
   override both.  A small search -- every problem of edit depth (budget plus
   the E13_FORCED cells: free cells and rules into the deleted state) at most
   E13_SMALL_DEPTH, about 1e8 leaves or fewer, which is bound by its longest
   machines rather than by throughput -- runs with the small profile unless
   settings were installed with e13_set_tuning since e13_open. ---- */
typedef struct {
  int32_t trials, pad;
  double first_ms, best_ms; /* wall time of the search call with the starting and the chosen settings */
} e13_tune_report;

int e13_autotune(e13_ctx *ctx, const e13_problem *ps, int np, int reps, e13_tuning *best, e13_tune_report *rep);
int e13_save_tuning(const e13_tuning *t, const char *path);
int e13_load_tuning(const char *path, e13_tuning *t);
int e13_tuning_file(e13_ctx *ctx, char *buf, int len);
int e13_tuning_file_small(e13_ctx *ctx, char *buf, int len);

#define E13_SMALL_DEPTH 3
enum { E13_PROFILE_MAIN = 0, E13_PROFILE_SMALL = 1 };
/* The settings profile a persistent search of ps[0..np) runs with (E13_PROFILE_MAIN or
   E13_PROFILE_SMALL); t (optional) receives its settings. */
int e13_search_profile(e13_ctx *ctx, const e13_problem *ps, int np, e13_tuning *t);

/* ---- Subtree cost model (no reference counterpart; the nearest is the Knuth
   estimator probe(), e13search.c:256-313, which estimates the same two
   quantities by random descents).  A 32 -> 64 -> 64 -> 2 perceptron (ReLU,
   fp32) maps a frontier node's features to (log1p(leaves)/16, log1p(steps)/24)
   of its subtree.  It only orders and splits the frontier of the persistent
   search: every subtree is still searched, so counts and survivors are the
   same with and without it. ---- */
#define E13_NFEAT 32
#define E13_MLP_PARAMS 6402 /* W1[64][32] b1[64] W2[64][64] b2[64] W3[2][64] b3[2] */

typedef struct {
  float x[E13_NFEAT];     /* features of the frontier node */
  uint64_t leaves, steps; /* of its subtree */
  int32_t prob, depth;
  double est_steps;       /* the subtree-steps estimate the frontier order used (learned = 1: the
                             network's, 3: the probes' mean; 1e300 for a node paused at a cap), else 0 */
} e13_cost_rec;

/* Frontier log of the last persistent search run with tuning.cost_log = 1. */
int e13_get_cost_log(e13_ctx *ctx, e13_cost_rec *out, int64_t cap, int64_t *n);
/* Regression targets of a subtree: t[0] = log1p(leaves)/16, t[1] = log1p(steps)/24. */
void e13_cost_targets(uint64_t leaves, uint64_t steps, float *t2);
/* Deterministic He-uniform initialisation from splitmix64(seed). */
void e13_mlp_init(uint64_t seed, float *params);
/* y[n][2] = f(x[n][32]) on the GPU. */
int e13_mlp_forward(e13_ctx *ctx, const float *params, const float *x, int64_t n, float *y);
/* Gradient of L = (1/n) sum (f(x) - t)^2 over all n samples (GPU), and L. */
int e13_mlp_grad(e13_ctx *ctx, const float *params, const float *x, const float *t, int64_t n, float *grad,
                 double *loss);
/* Adam training on the GPU: `epochs` passes over a seeded shuffle in batches
   of `batch`; params in/out; loss_per_epoch[epochs] (may be NULL) holds the
   mean loss over each epoch's batches. */
int e13_mlp_train(e13_ctx *ctx, float *params, const float *x, const float *t, int64_t n, int epochs, int batch,
                  float lr, uint64_t seed, float *loss_per_epoch);
/* Install (or with NULL remove) the model used when tuning.learned = 1. */
int e13_set_cost_model(e13_ctx *ctx, const float *params);

#ifdef __cplusplus
}
#endif
#endif
