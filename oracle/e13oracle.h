/* e13oracle.h -- CPU oracle for the TNF enumeration + simulate-and-filter path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C11/OpenMP restatement of the
 * reference search (src/e13search.c) used to check the HIP implementation and
 * to time the CPU baseline.  Only tests, the smoke check and the CPU-baseline
 * leg of bench.py may load or run it; the product path never does.
 *
 * Every function names the reference lines it restates.  Pinning (what checks
 * each function against something other than itself) is listed in the test
 * files: golden fixtures from the reference's results/ tree, dumps of the
 * compiled reference, closed forms and a brute-force set simulator.
 */
#ifndef E13ORACLE_H
#define E13ORACLE_H

#include <stdint.h>

#define EO_NS 9                 /* stages 1..9 (e13search.c:58) */
#define EO_LO 64                /* cells left of 0 (e13search.c:59) */
#define EO_HI 1300              /* cells from 0 rightwards (e13search.c:60) */
#define EO_TAPE (EO_LO + EO_HI) /* 1364 (e13search.c:61) */
#define EO_MAXQ 16              /* e13search.c:63 */
#define EO_DEAD (-1)            /* run() result (e13search.c:64) */
#define EO_SURV (-2)            /* run() result (e13search.c:65) */
#define EO_FREE (-2)            /* table code of a cell still to be chosen (FORCED, e13search.c:66) */
#define EO_HALT (-1)            /* table code of a halt */

#ifdef __cplusplus
extern "C" {
#endif

/* Zobrist keys KEY[0..1363] and stage targets TGT[1..9] (e13search.c:68-69, 437-444). */
extern uint64_t eo_key[EO_TAPE];
extern uint64_t eo_tgt[EO_NS + 1];
void eo_init_keys(void);
uint64_t eo_splitmix(uint64_t *state); /* e13search.c:98-103 */

/* Step at which erdos_14 reaches stage n, and the counter class of stage n
   (e13search.c:70-71). */
extern const long eo_t14s[EO_NS + 1];
extern const int eo_cval[EO_NS + 1];

/* Machine state (e13search.c:77-82). */
typedef struct {
  int head, st, nt, stage, st0, head0, abs_ok, rel_ok;
  int cs[3];
  long long t, deadline;
  uint64_t hb, hc, ch[3];
  uint8_t tape[EO_TAPE];
} eo_sim;

/* Per-stage deadlines dl[1..9]; NULL means the reference's 4*T14S[n]+64. */
void eo_sim_init(eo_sim *s, const int32_t *dl);                                  /* e13search.c:123-129 */
int eo_run(eo_sim *s, const int *tab, const uint8_t *fixed, const int32_t *dl); /* e13search.c:131-177 */

/* Parse one cell "1R5" / "---" into a code (e13search.c:105-108). */
int eo_parse_cell(const char *s);
/* Load a machine file: returns the number of states, codes into base[2*MAXQ]
   (e13search.c:110-121); -1 when the file cannot be opened. */
int eo_load(const char *path, int *base);

/* Search problem after set_sym()/setup() (e13search.c:315-355, 422-426). */
typedef struct {
  int nq_base;              /* states in the loaded machine */
  int base[2 * EO_MAXQ];    /* loaded machine */
  int X;                    /* deleted state, -1 = none */
  int fmax;                 /* delete mode: edit budget; free mode: ignored */
  int freemode;             /* 1: free/freesym (budget 0), 0: delete */
  uint8_t freebase[2 * EO_MAXQ];
  int sym_lo, sym_hi;       /* renumbered, -1 = off */
  /* filled by eo_setup */
  int nq, budget, sym_n0;
  int tab[2 * EO_MAXQ];
} eo_problem;

/* Returns 0, or -1 when the forced cells exceed FMAX (budget < 0), or -2 for a
   symmetric-state violation, -3 for the start state inside the symmetric set. */
int eo_setup(eo_problem *p);
void eo_set_sym(eo_problem *p, int lo, int hi); /* e13search.c:423-426 */

typedef struct {
  long long leaves, survivors, steps;
  long long hist[EO_NS + 1];
  long nfront;
} eo_counts;

typedef void (*eo_surv_fn)(void *user, const eo_sim *s, const int *tab, const uint8_t *fixed, int X, int nq);

/* run_delete(): serial front to `split` branchings, then OpenMP dynamic(1)
   over the front with replay from the blank tape (e13search.c:357-420).
   `threads` <= 0 uses the OpenMP default.  Survivors are passed to fn
   under a critical section. */
int eo_search(eo_problem *p, int split, int threads, eo_counts *out, eo_surv_fn fn, void *user);

/* Full-table run from the blank tape (run_full, e13search.c:428-433), with
   optional per-stage deadlines.  Returns EO_DEAD or EO_SURV. */
int eo_run_full(const int *tab, eo_sim *s, const int32_t *dl);

/* Plain, unhashed re-check of a survivor restricted to stages 1..9 with the
   search's own deadlines (after recheck.c:57-92): exact block, exact counter
   sets, no hashing.  Returns 1 if consistent, 0 if not; the read mask of the
   cells touched before the stage-9 start goes to *read_mask. */
int eo_recheck(const int *tab, int nq, int st0, int head0, int abs_ok, int rel_ok, long long steps,
               uint32_t *read_mask);

/* Synthetic batching stress table k (28 cells of a 14-state machine, one halt
   cell, log-uniform step cap in [10, 2e5]); see e13oracle_core.c. */
void eo_synth_table(uint64_t seed0, uint64_t k, int *tab28, int32_t *cap);

/* Order-independent digest and totals of the synthetic tables k0..k0+n-1, each
   run with its cap as every stage deadline (see e13oracle_core.c for the
   record hash). */
typedef struct {
  uint64_t n, digest, steps, survivors;
  uint64_t hist[EO_NS + 1];
} eo_synth_sum;
uint64_t eo_synth_record_hash(uint64_t k, int result, int stage, long long t, int st0, int head0, int abs_ok,
                              int rel_ok, const uint64_t ch[3]);
void eo_synth_digest(uint64_t seed0, uint64_t k0, long long n, int threads, eo_synth_sum *out);

/* Probe estimator (probe(), e13search.c:256-313; estimate mode, :473-510).
   Probe p draws from its own stream, state eo_probe_stream(seed, p); the
   per-probe sums are leaves, nodes, steps, survivors (weighted). */
uint64_t eo_probe_stream(uint64_t seed, uint64_t p);
void eo_probe(const eo_problem *p, uint64_t *rng, double out[4]);
typedef struct {
  long long probes;
  double sum[4], sq[4], mean[4], se[4]; /* leaves, nodes, steps, survivors */
} eo_estimate_sum;
int eo_estimate(const eo_problem *p, uint64_t seed, long long probes, int threads, double *per,
                eo_estimate_sum *out);

#ifdef __cplusplus
}
#endif
#endif
