/* e13_ctx.h -- private to libe13.so: the device context shared by the host
 * C files, the launch wrappers of the HIP files, and internal helpers. */
#ifndef E13_CTX_H
#define E13_CTX_H
#define __HIP_PLATFORM_AMD__ 1
#include <hip/hip_runtime_api.h>
#include <stdio.h>

#include "e13.h"
#include "e13_dev.h"

#define HIPCHK(x)                                                                                                     \
  do {                                                                                                                \
    hipError_t e_ = (x);                                                                                              \
    if (e_ != hipSuccess) {                                                                                           \
      fprintf(stderr, "libe13: %s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), __FILE__, __LINE__);             \
      return E13_EHIP;                                                                                                \
    }                                                                                                                 \
  } while (0)

/* Launch wrappers (e13_kernels.hip). */
size_t e13k_lds_bytes(int wg_threads, int gate);
int e13k_launch_expand(const E13ExpandArgs *a, int grid, int wg_threads, int gate, int refill, hipStream_t s);
int e13k_launch_batch(const E13BatchArgs *a, int grid, int wg_threads, int gate, int refill, hipStream_t s);
size_t e13k_search_lds_bytes(int wg_threads, int gate, int np, int tw);
int e13k_search_max_wgs(int wg_threads, int gate, int np, int tw);
int e13k_launch_search(const E13SearchArgs *a, int grid, int wg_threads, int gate, hipStream_t s);
int e13k_launch_chain_pop(E13Chain *ch, E13Counters *ctr, hipStream_t s);
int e13k_launch_chain_push(E13Chain *ch, E13Node *stack, uint64_t stack_cap, const E13Node *out, uint32_t out_cap,
                           const E13Counters *ctr, int grid, hipStream_t s);

size_t e13k_probe_lds_bytes(int wg_threads, int gate);
int e13k_probe_max_wgs(int wg_threads, int gate);
int e13k_launch_probe(const E13ProbeArgs *a, int grid, int wg_threads, int gate, hipStream_t s);
int e13k_scratch_addr(int device, void **ptr);
#define E13_FRONTIER_PROBES 16 /* probes per frontier node when the frontier is ordered by probes */
void e13i_probe_root(const e13_problem *p, E13Node *nd);
int e13i_probe_nodes(e13_ctx *c, const e13_problem *p, const E13Problem *pb_dev, const E13Node *nodes, size_t n,
                     int per_node, uint64_t seed, double *est);
int e13k_launch_zero(void *p, uint32_t nwords, hipStream_t s);

_Static_assert(sizeof(E13Node) % 16 == 0, "node records are copied as 16-byte vectors");


struct e13_ctx {
  int device, ncu;
  char arch[256]; /* gcnArchName */
  size_t lds_max; /* largest dynamic LDS a workgroup may use */
  size_t lds_cu;  /* LDS of one CU */
  int narrow_off; /* a short-tape persistent launch found workgroups not resident: whole tape only */
  hipStream_t stream;
  hipEvent_t ev0, ev1;
  e13_tuning tune;
  e13_tuning tune_small; /* settings for small searches (valid when has_small) */
  e13_tuning tune_open, tune_small_open; /* both profiles as e13_open installed them */
  int has_small, auto_profile; /* auto_profile: no e13_set_tuning since e13_open */
  e13_stats stats;
  E13Counters *d_ctr;
  uint64_t *d_key, *d_tgt;
  E13Problem *d_pb;
  E13Node *d_proot;      /* the probe estimator's root node */
  E13Counters *d_pctr;   /* counters of the frontier's probe launches (the search keeps d_ctr) */
  double *seed_est;      /* subtree-steps estimate of each frontier node, in frontier order (the
                            order's cost: learned = 1 the network, 3 the probes), for the cost log */
  size_t seed_est_n, seed_est_cap;
  int32_t *d_dl;
  E13Node *d_stack;
  size_t stack_cap;
  E13Node *d_out;
  size_t out_cap;
  E13SurvRec *d_surv;
  size_t surv_cap;
  void *d_tabs, *d_caps, *d_leaves;
  size_t batch_cap;
  /* persistent search */
  E13Node *d_stacks;
  size_t stacks_cap; /* nodes */
  E13Mail *d_mail;     /* one mailbox per wave */
  uint64_t *d_hungry;  /* hungry map */
  size_t mail_cap;     /* waves */
  E13Sched *d_sch;
  E13Problem *d_pbs;
  uint64_t *d_pcnt;
  E13Node *d_seeds;
  size_t seeds_cap;
  E13Chain *d_chain; /* launch-chain cursor of the host-loop forms */
  uint64_t *d_trace; /* wave event log of the last traced search */
  size_t trace_words;
  int trace_waves, trace_per_wave;
  /* cost model (e13_cost.c) */
  float *d_mlp;       /* params, grad, Adam m, Adam v: 4 x E13_MLP_PARAMS */
  float *d_mlp_x, *d_mlp_t, *d_mlp_y, *d_mlp_part, *d_mlp_sse;
  int32_t *d_mlp_idx;
  size_t mlp_rows;    /* rows the sample buffers hold */
  int have_model;     /* a cost model is installed (e13_set_cost_model) */
  float model[E13_MLP_PARAMS];
  uint64_t *d_seed_cnt;
  size_t seed_cnt_cap;
  e13_cost_rec *cost_log; /* per frontier node of the last logged search */
  int64_t cost_n, cost_cap;
  /* probe estimator (e13_estimate.c) */
  double *d_probe_out;
  size_t probe_cap;      /* probes the output buffer holds */
  void *scratch;         /* this device's static scratch (e13_estimate_static), NULL when not found */
  /* checkpoints of the host-loop search (e13_ckpt.c) */
  char ckpt_path[1024];  /* empty: off */
  double ckpt_every_s;
  int64_t ckpt_max_launches;
  int ckpt_slot;         /* 0: the next save goes to <path>.a, 1: to <path>.b */
  uint32_t *h_stop, *d_stop; /* the persistent form's clean-stop flag: pinned host memory, its device address */
  int64_t *d_wave_state;     /* per wave: stack range and last delivery tag, written when a wave leaves */
  size_t wave_state_cap;
  E13Node *d_wide;           /* the narrow form's list of machines past its tape, for the whole-tape launch */
  size_t wide_cap;
};

/* The MLP kernels (e13_mlp.hip). */
int e13k_mlp_forward(const float *P, const float *X, const int32_t *idx, int n, float *Y, hipStream_t s);
int e13k_mlp_grad(const float *P, const float *X, const float *T, const int32_t *idx, int row0, int m, float inv_n,
                  float *part, float *sse, float *grad, hipStream_t s);
int e13k_mlp_adam(float *P, const float *G, float *M, float *V, float lr, float b1, float b2, float eps, float c1,
                  float c2, hipStream_t s);
int e13k_launch_seed_ids(E13Node *seeds, uint32_t n, hipStream_t s);

/* Internal helpers (e13_host.c, e13_cost.c); hidden from the ABI. */
#define E13_INTERNAL __attribute__((visibility("hidden")))
E13_INTERNAL int e13i_ensure_buf(void **p, size_t *cap, size_t need, size_t elem);
E13_INTERNAL uint64_t e13i_splitmix(uint64_t *s);
E13_INTERNAL void e13i_root_node(const e13_problem *p, E13Node *nd);
E13_INTERNAL void e13i_fill_problem(const e13_problem *p, E13Problem *pb);
/* cost model: features of frontier nodes; the learned order of a frontier */
E13_INTERNAL void e13i_node_features(const e13_problem *p, const E13Node *nd, float *f);
E13_INTERNAL int e13i_mlp_predict_dev(e13_ctx *c, const float *x_host, int n, float *y_host);
E13_INTERNAL void e13i_free_cost(e13_ctx *c);
E13_INTERNAL void e13i_free_estimate(e13_ctx *c);
E13_INTERNAL int e13i_ckpt_save(e13_ctx *c, const e13_problem *p, size_t top, size_t peak, int64_t launches,
                                double wall_ms);
E13_INTERNAL int e13i_ckpt_stop_requested(e13_ctx *c);
E13_INTERNAL void e13i_ckpt_stop_clear(e13_ctx *c);
E13_INTERNAL int e13i_ckpt3_prepare(e13_ctx *c, int total_waves);
E13_INTERNAL int e13i_ckpt3_gather(e13_ctx *c, int total_waves, int stk_cap, const E13Node *extra, size_t n_extra,
                                  size_t *n_left, size_t *nseeds);
E13_INTERNAL int e13i_ckpt3_save(e13_ctx *c, const e13_problem *ps, int np, const e13_counts *out, size_t nseeds);
E13_INTERNAL int e13i_ckpt3_load(e13_ctx *c, const e13_problem *ps, int np, e13_counts *out, size_t *nseeds);
E13_INTERNAL int e13i_ckpt_load(e13_ctx *c, const e13_problem *p, size_t *top, size_t *peak, int64_t *launches,
                                double *wall_ms);
/* stored per-device tuning (e13_tune.c) */
E13_INTERNAL int e13i_tuning_path(const char *arch, int ncu, int small, char *buf, int len);

#endif
