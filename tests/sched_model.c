/* sched_model.c -- a CPU model of the work-sharing protocol of the persistent
 * search kernel (k_search in erdos-machine-13_amd/csrc/e13_kernels.hip): host
 * threads stand in for waves, with random pauses between every shared-memory
 * operation so that runs see many interleavings.
 *
 * What is modelled, in the kernel's order per loop iteration:
 *   - private node stacks (rings of stk_cap nodes) with a staged top node and
 *     its child-slot cursor [cur_slot, cur_end), saved into the node when the
 *     wave stages another one or gives the node away;
 *   - refill of idle lanes from the staged node's slots;
 *   - the mailbox poll of a wave that asked for work: the flag, then the
 *     elements re-read until every one carries the flag's tag;
 *   - the request (one bit in the hungry map) once the stack is empty and
 *     enough lanes are idle, and its withdrawal when work of its own appears;
 *   - the busy count: waves holding work plus deliveries in flight; the wave
 *     whose decrement takes it to zero writes the end mark into every mailbox;
 *   - the idle watchdog, the stack-full and lost-delivery errors;
 *   - lanes that run for a random number of rounds and stop as a leaf or as a
 *     node with child slots; eviction of long runners as one-slot resumable
 *     nodes when hungry waves exist;
 *   - donation: the oldest node, or the upper part of the staged node's slots,
 *     to a hungry wave won by clearing its bit; the receiver's delivery number
 *     from its mailbox's sequence counter, the tag seq % (2^bits - 1) + 1.
 * The tree is a hash-defined random tree; the run's leaf, machine and step
 * totals must equal a sequential walk of the same tree, and every wave must
 * leave through the end mark.
 *
 * Planted mistakes (each must make some run fail):
 *   --plant-empty     launch with no seeds and no early return (no wave ever
 *                     counts busy down: the watchdog ends it);
 *   --plant-donor-tag a tag made of the donor's id and its give count modulo
 *                     2^bits (a receiver can see its last tag again and skip a
 *                     delivery the donor counted as busy);
 *   --plant-nobusy    a delivery not counted as busy (the end mark can go out
 *                     while a node is in flight: totals come up short).
 *
 * Output: one line, "ok leaves machines steps gives takes" or "error K ...".
 * Exit 0 when the run ends cleanly with the walk's totals, 1 otherwise.
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <sched.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define LANES 8
#define NODE_WORDS 6 /* id, depth, resume, nslot, next_slot, check */
#define MAIL_DONE UINT64_MAX
#define MAXW 256

typedef struct {
  uint64_t id;
  uint64_t depth, resume, nslot, next_slot, check;
} node;

typedef struct {
  _Atomic uint64_t flag;
  _Atomic uint64_t seq;
  _Atomic uint64_t w[NODE_WORDS];
} mailbox;

/* ---- parameters ---- */
static int W = 8, map_bits = 64, stk_cap = 1024, tag_bits = 32, depth_max = 6, max_slots = 5, nseeds = 16;
static int pool_min_idle = 4, donate_every = 2, sched_every = 4, evict_after = 40, refill_tries = 2;
static int plant_empty, plant_donor_tag, plant_nobusy, monitor;
static double idle_limit_s = 3.0, flag_first_p = 0.2;
static uint64_t tree_seed = 1;
static double spin_limit_s = 2.0; /* lost delivery: elements not all tagged after this long */

/* ---- shared state ---- */
static mailbox mail[MAXW];
static _Atomic uint64_t hungry[MAXW];
static _Atomic int busy_count, sched_error;
static _Atomic uint64_t tot_leaves, tot_machines, tot_steps, tot_gives, tot_takes, exited;
static node *seeds;
static _Atomic long long st_top[MAXW], st_bot[MAXW], st_act[MAXW], st_req[MAXW], st_busy[MAXW], st_iter[MAXW];

/* ---- the tree ---- */
static uint64_t mix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
static uint64_t h2(uint64_t a, uint64_t b) { return mix(a ^ mix(b + tree_seed)); }
static int nslots_of(uint64_t id) { return 1 + (int)(h2(id, 1) % (uint64_t)max_slots); }
static int slot_ok(uint64_t id, int s) { return h2(id, 100 + (uint64_t)s) % 4 != 0; }
static uint64_t child_of(uint64_t id, int s) { return h2(id, 200 + (uint64_t)s); }
static int run_len(uint64_t m) { /* mostly short, a few long runners */
  const uint64_t r = h2(m, 3);
  return r % 40 == 0 ? 200 + (int)((r >> 8) % 400) : 1 + (int)((r >> 8) % 30);
}
static int stops_live(uint64_t m, int depth) { return depth < depth_max && h2(m, 4) % 5 < 3; }

static void walk(uint64_t id, int depth, int nslot, uint64_t *lv, uint64_t *mc, uint64_t *st) {
  for (int s = 0; s < nslot; ++s) {
    if (!slot_ok(id, s)) continue;
    const uint64_t m = child_of(id, s);
    *mc += 1;
    *st += (uint64_t)run_len(m);
    if (stops_live(m, depth + 1)) walk(m, depth + 1, nslots_of(m), lv, mc, st);
    else *lv += 1;
  }
}

/* ---- per-thread randomness and pauses ---- */
typedef struct {
  uint64_t s;
} rng;
static uint64_t rnext(rng *r) {
  r->s ^= r->s << 13;
  r->s ^= r->s >> 7;
  r->s ^= r->s << 17;
  return r->s;
}
static void pause_at(rng *r) {
  const uint64_t v = rnext(r) & 63;
  if (v == 0) sched_yield();
  else if (v < 8)
    for (volatile int i = 0; i < (int)v * 40; ++i) {
    }
}
static double now_s(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static uint64_t node_check(const node *n) { return mix(n->id ^ (n->depth << 7) ^ (n->resume << 17) ^ (n->nslot << 37)); }

static void broadcast_done(void) {
  for (int w = 0; w < W; ++w) atomic_store(&mail[w].flag, MAIL_DONE);
}

typedef struct {
  int active;
  uint64_t m;
  int depth, left, age;
} lane_t;

static void *wave(void *arg) {
  const int wid = (int)(intptr_t)arg;
  rng R = {mix((uint64_t)wid + 77 * tree_seed) | 1};
  node *stk = calloc((size_t)stk_cap, sizeof(node));
  lane_t L[LANES];
  memset(L, 0, sizeof L);
  long long top = 0, bot = 0, staged = -1;
  int cur_slot = 0, cur_end = 0, busy = 0, requested = 0, mail_in = 1, refresh_in = wid % 8 + 1;
  node stg; /* the staged node */
  memset(&stg, 0, sizeof stg);
  uint64_t last_tag = 0, gives = 0, takes = 0, leaves = 0, machines = 0, steps = 0;
  double idle_since = 0;
  int idle_polls = 0;
  const int words = (W + map_bits - 1) / map_bits;
  const uint64_t wmask = map_bits == 64 ? ~0ull : (1ull << map_bits) - 1;
#define NODE(i) stk[(size_t)((i) & (stk_cap - 1))]
  /* seeds: seed i goes to wave i mod W (the host counted each such wave busy) */
  if (!plant_empty && nseeds == 0) goto out; /* no work: no wave would ever count busy down */
  {
    const int mine = nseeds / W + (wid < nseeds % W ? 1 : 0);
    for (int j = 0; j < mine; ++j) NODE(j) = seeds[wid + j * W];
    top = mine;
    busy = mine > 0;
  }
  for (;;) {
    /* ---- refill idle lanes from the staged node ---- */
    for (int tries = 0; tries < refill_tries; ++tries) {
      int need = 0;
      for (int l = 0; l < LANES; ++l) need += !L[l].active;
      if (need == 0 || top == bot) break;
      const long long tidx = top - 1;
      if (staged != tidx) {
        if (staged >= bot && staged < top) NODE(staged).next_slot = (uint64_t)cur_slot | ((uint64_t)cur_end << 16);
        stg = NODE(tidx);
        staged = tidx;
        cur_slot = (int)(stg.next_slot & 0xFFFF);
        cur_end = (int)(stg.next_slot >> 16);
        if (cur_end == 0) cur_end = (int)stg.nslot;
      }
      const int give = need < cur_end - cur_slot ? need : cur_end - cur_slot;
      int r = 0;
      for (int l = 0; l < LANES && r < give; ++l) {
        if (L[l].active) continue;
        const int slot = cur_slot + r++;
        if (stg.resume) { /* a paused machine: the node is its own single child */
          L[l] = (lane_t){1, stg.id, (int)stg.depth, (int)stg.resume, 0};
        } else if (slot_ok(stg.id, slot)) {
          const uint64_t m = child_of(stg.id, slot);
          L[l] = (lane_t){1, m, (int)stg.depth + 1, run_len(m), 0};
          machines++;
        }
      }
      cur_slot += give;
      if (cur_slot >= cur_end) {
        top -= 1;
        staged = -1;
      }
    }
    /* ---- the mailbox of a wave that asked for work ---- */
    if (requested && --mail_in <= 0) {
      mail_in = 1;
      pause_at(&R);
      const uint64_t f = atomic_load(&mail[wid].flag);
      if (f == MAIL_DONE) break;
      if (f != last_tag) {
        if (top - bot + 1 > stk_cap) {
          int z = 0;
          atomic_compare_exchange_strong(&sched_error, &z, 1);
          broadcast_done();
          break;
        }
        node n;
        uint64_t e[NODE_WORDS];
        int ok = 0;
        const double t0 = now_s();
        for (;;) {
          ok = 1;
          for (int k = 0; k < NODE_WORDS; ++k) {
            e[k] = atomic_load(&mail[wid].w[k]);
            ok &= (uint32_t)(e[k] >> 32) == (uint32_t)f;
          }
          if (ok || now_s() - t0 > spin_limit_s) break;
          pause_at(&R);
        }
        if (!ok) {
          int z = 0;
          atomic_compare_exchange_strong(&sched_error, &z, 2);
          broadcast_done();
          break;
        }
        uint64_t low[NODE_WORDS]; /* the low 32 bits of each element */
        for (int k = 0; k < NODE_WORDS; ++k) low[k] = (uint32_t)e[k];
        n.id = low[0] | ((uint64_t)low[1] << 32);
        n.depth = low[2] & 0xFF;
        n.resume = low[2] >> 8;
        n.nslot = low[3];
        n.next_slot = low[4];
        n.check = low[5];
        if ((uint32_t)node_check(&n) != (uint32_t)n.check) { /* a torn delivery */
          int z = 0;
          atomic_compare_exchange_strong(&sched_error, &z, 5);
          broadcast_done();
          break;
        }
        NODE(top) = n;
        if (busy) atomic_fetch_sub(&busy_count, 1); /* the donor's count returned */
        last_tag = f;
        requested = 0;
        busy = 1;
        idle_since = 0;
        top += 1;
        ++takes;
        continue;
      }
    }
    /* ---- ask for work ---- */
    if (!requested && top == bot) {
      int idle = 0, long_runner = 0;
      for (int l = 0; l < LANES; ++l) {
        idle += !L[l].active;
        long_runner |= L[l].active && L[l].age >= evict_after;
      }
      const int hold = evict_after > 0 && idle == LANES - 1 && long_runner;
      if ((idle >= pool_min_idle || idle == LANES) && !hold) {
        pause_at(&R);
        atomic_fetch_or(&hungry[wid / map_bits], 1ull << (wid % map_bits));
        requested = 1;
        mail_in = 1;
      }
    }
    int act = 0;
    for (int l = 0; l < LANES; ++l) act += L[l].active;
    if (monitor) {
      atomic_store(&st_top[wid], top);
      atomic_store(&st_bot[wid], bot);
      atomic_store(&st_act[wid], act);
      atomic_store(&st_req[wid], requested);
      atomic_store(&st_busy[wid], busy);
      atomic_fetch_add(&st_iter[wid], 1);
    }
    if (act == 0) {
      if (top != bot) continue;
      if (busy) {
        pause_at(&R);
        const int last = atomic_fetch_sub(&busy_count, 1) == 1;
        busy = 0;
        if (last) {
          broadcast_done();
          break;
        }
      }
      const double t = now_s();
      if (idle_since == 0) idle_since = t;
      if ((++idle_polls & 63) == 0 && atomic_load(&sched_error)) break;
      if (t - idle_since > idle_limit_s) {
        int z = 0;
        atomic_compare_exchange_strong(&sched_error, &z, 3);
        broadcast_done();
        break;
      }
      sched_yield();
      continue;
    }
    idle_since = 0;
    const int refresh = --refresh_in <= 0;
    uint64_t hw[MAXW];
    int s_err = 0, any_hungry = 0;
    if (refresh) {
      s_err = atomic_load(&sched_error);
      for (int k = 0; k < words; ++k) {
        hw[k] = atomic_load(&hungry[k]) & wmask;
        any_hungry |= hw[k] != 0;
      }
    }
    /* ---- one round of the lanes ---- */
    int ev[LANES] = {0}; /* 1 leaf, 2 stop with children, 3 paused (cap) */
    for (int l = 0; l < LANES; ++l) {
      if (!L[l].active) continue;
      L[l].left -= 1;
      L[l].age += 1;
      steps += 1;
      if (L[l].left == 0) ev[l] = stops_live(L[l].m, L[l].depth) ? 2 : 1;
    }
    /* ---- long runners to hungry waves, paused as one-slot nodes ---- */
    if (refresh && evict_after > 0 && top == bot && any_hungry) {
      int running = 0, n_ev = 0;
      for (int l = 0; l < LANES; ++l) running += L[l].active && !ev[l];
      for (int l = 0; l < LANES && n_ev < donate_every && n_ev < running - 1; ++l)
        if (L[l].active && !ev[l] && L[l].age >= evict_after) {
          ev[l] = 3;
          ++n_ev;
        }
    }
    /* ---- finished lanes; live ones onto the stack ---- */
    int n_emit = 0;
    for (int l = 0; l < LANES; ++l) n_emit += ev[l] >= 2;
    if (n_emit) {
      if (top - bot + n_emit > stk_cap) {
        int z = 0;
        atomic_compare_exchange_strong(&sched_error, &z, 1);
        broadcast_done();
        break;
      }
      if (staged >= bot && staged < top) NODE(staged).next_slot = (uint64_t)cur_slot | ((uint64_t)cur_end << 16);
    }
    for (int l = 0; l < LANES; ++l) {
      if (!ev[l]) continue;
      if (ev[l] == 1) leaves++;
      else {
        node n = {L[l].m, (uint64_t)L[l].depth, ev[l] == 3 ? (uint64_t)L[l].left : 0,
                  ev[l] == 3 ? 1 : (uint64_t)nslots_of(L[l].m), 0, 0};
        n.check = (uint32_t)node_check(&n);
        NODE(top) = n;
        top += 1;
      }
      L[l].active = 0;
    }
    if (n_emit) { /* the newest node is staged at once */
      stg = NODE(top - 1);
      staged = top - 1;
      cur_slot = 0;
      cur_end = (int)stg.nslot;
    }
    /* ---- work of its own again: withdraw the request unless a donor took it ---- */
    if (requested && top > bot) {
      pause_at(&R);
      const uint64_t bit = 1ull << (wid % map_bits);
      if (atomic_fetch_and(&hungry[wid / map_bits], ~bit) & bit) requested = 0;
    }
    if (!refresh) continue;
    if (s_err) break;
    /* ---- deliveries to hungry waves ---- */
    int given = 0, idle_now = 0;
    for (int l = 0; l < LANES; ++l) idle_now += !L[l].active;
    for (;;) {
      int active_now = 0;
      for (int l = 0; l < LANES; ++l) active_now += L[l].active;
      const int whole = top - bot >= 2 || (top - bot == 1 && active_now && evict_after > 0);
      const int rem = cur_end - cur_slot, keep = idle_now > rem / 2 ? idle_now : rem / 2;
      const int split = !whole && top - bot == 1 && staged == top - 1 && rem >= 2 && keep < rem;
      int wsel = -1;
      for (int k = 0; k < words && wsel < 0; ++k) {
        const int kk = (k + wid) % words; /* donors spread over the map */
        if (hw[kk]) wsel = kk;
      }
      if (wsel < 0 || given >= donate_every || !(whole || split)) break;
      const int brot = (wid * 13 + given) % map_bits;
      int bsel = -1;
      for (int b = 0; b < map_bits; ++b) {
        const int bb = (b + brot) % map_bits;
        if ((hw[wsel] >> bb) & 1) {
          bsel = bb;
          break;
        }
      }
      const uint64_t bit = 1ull << bsel;
      hw[wsel] &= ~bit;
      const int dst = wsel * map_bits + bsel;
      pause_at(&R);
      if (!(atomic_fetch_and(&hungry[wsel], ~bit) & bit)) continue; /* another donor won it */
      if (!plant_nobusy) atomic_fetch_add(&busy_count, 1);         /* in flight counts as busy */
      const uint64_t dseq = atomic_fetch_add(&mail[dst].seq, 1);
      node n;
      if (whole) {
        if (staged == bot) {
          NODE(staged).next_slot = (uint64_t)cur_slot | ((uint64_t)cur_end << 16);
          staged = -1;
        }
        n = NODE(bot);
      } else {
        n = stg;
        n.next_slot = (uint64_t)(cur_slot + keep) | ((uint64_t)cur_end << 16);
        cur_end = cur_slot + keep;
      }
      n.check = (uint32_t)node_check(&n);
      uint64_t tag;
      if (plant_donor_tag) tag = ((uint64_t)(wid + 1) << tag_bits) | (gives & ((1ull << tag_bits) - 1));
      else tag = dseq % ((1ull << tag_bits) - 1) + 1;
      const uint64_t low[NODE_WORDS] = {(uint32_t)n.id, n.id >> 32, n.depth | (n.resume << 8), n.nslot, n.next_slot,
                                        n.check};
      /* elements in a random order; now and then the flag is seen first */
      const int flag_first = (double)(rnext(&R) % 1000) < flag_first_p * 1000.0;
      if (flag_first) atomic_store(&mail[dst].flag, tag);
      int order[NODE_WORDS];
      for (int k = 0; k < NODE_WORDS; ++k) order[k] = k;
      for (int k = NODE_WORDS - 1; k > 0; --k) {
        const int j = (int)(rnext(&R) % (uint64_t)(k + 1)), t = order[k];
        order[k] = order[j];
        order[j] = t;
      }
      for (int k = 0; k < NODE_WORDS; ++k) {
        atomic_store(&mail[dst].w[order[k]], (tag << 32) | low[order[k]]);
        pause_at(&R);
      }
      if (!flag_first) atomic_store(&mail[dst].flag, tag);
      if (whole) bot += 1;
      ++given;
      ++gives;
    }
    int nz = 0;
    for (int k = 0; k < words; ++k) nz |= hw[k] != 0;
    refresh_in = given > 0 || nz ? 2 : sched_every;
  }
out:
  atomic_fetch_add(&tot_leaves, leaves);
  atomic_fetch_add(&tot_machines, machines);
  atomic_fetch_add(&tot_steps, steps);
  atomic_fetch_add(&tot_gives, gives);
  atomic_fetch_add(&tot_takes, takes);
  atomic_fetch_add(&exited, 1);
  free(stk);
  return NULL;
#undef NODE
}

int main(int argc, char **argv) {
  for (int i = 1; i < argc; ++i) {
    const char *a = argv[i];
    const char *v = i + 1 < argc ? argv[i + 1] : "0";
    if (!strcmp(a, "--waves")) W = atoi(v), ++i;
    else if (!strcmp(a, "--map-bits")) map_bits = atoi(v), ++i;
    else if (!strcmp(a, "--stack")) stk_cap = atoi(v), ++i;
    else if (!strcmp(a, "--tag-bits")) tag_bits = atoi(v), ++i;
    else if (!strcmp(a, "--depth")) depth_max = atoi(v), ++i;
    else if (!strcmp(a, "--slots")) max_slots = atoi(v), ++i;
    else if (!strcmp(a, "--seeds")) nseeds = atoi(v), ++i;
    else if (!strcmp(a, "--tree")) tree_seed = strtoull(v, NULL, 10), ++i;
    else if (!strcmp(a, "--idle-limit")) idle_limit_s = atof(v), ++i;
    else if (!strcmp(a, "--evict")) evict_after = atoi(v), ++i;
    else if (!strcmp(a, "--pool-min-idle")) pool_min_idle = atoi(v), ++i;
    else if (!strcmp(a, "--flag-first")) flag_first_p = atof(v), ++i;
    else if (!strcmp(a, "--plant-empty")) plant_empty = 1;
    else if (!strcmp(a, "--monitor")) monitor = 1;
    else if (!strcmp(a, "--plant-donor-tag")) plant_donor_tag = 1;
    else if (!strcmp(a, "--plant-nobusy")) plant_nobusy = 1;
    else {
      fprintf(stderr, "unknown option %s\n", a);
      return 2;
    }
  }
  if (W < 1 || W > MAXW || map_bits < 1 || map_bits > 64 || (stk_cap & (stk_cap - 1)) || tag_bits < 1 ||
      tag_bits > 32 || W > map_bits * MAXW) {
    fprintf(stderr, "bad parameters\n");
    return 2;
  }
  /* the frontier: seed nodes at depth 0 */
  seeds = calloc((size_t)(nseeds > 0 ? nseeds : 1), sizeof(node));
  uint64_t want_lv = 0, want_mc = 0, want_st = 0;
  for (int i = 0; i < nseeds; ++i) {
    const uint64_t id = h2((uint64_t)i, 9);
    seeds[i] = (node){id, 0, 0, (uint64_t)nslots_of(id), 0, 0};
    seeds[i].check = (uint32_t)node_check(&seeds[i]);
    walk(id, 0, nslots_of(id), &want_lv, &want_mc, &want_st);
  }
  int holders = 0;
  for (int w = 0; w < W; ++w) holders += nseeds / W + (w < nseeds % W ? 1 : 0) > 0;
  atomic_store(&busy_count, holders);
  pthread_t th[MAXW];
  for (int w = 0; w < W; ++w) pthread_create(&th[w], NULL, wave, (void *)(intptr_t)w);
  if (monitor)
    while ((int)atomic_load(&exited) < W) {
      struct timespec ts = {1, 0};
      nanosleep(&ts, NULL);
      fprintf(stderr, "busy %d err %d |", atomic_load(&busy_count), atomic_load(&sched_error));
      for (int w = 0; w < W; ++w)
        fprintf(stderr, " %d:%lld/%lld a%lld r%lld b%lld i%lld", w, atomic_load(&st_top[w]) - atomic_load(&st_bot[w]),
                atomic_load(&st_top[w]), atomic_load(&st_act[w]), atomic_load(&st_req[w]), atomic_load(&st_busy[w]),
                atomic_load(&st_iter[w]));
      fprintf(stderr, "\n");
    }
  for (int w = 0; w < W; ++w) pthread_join(th[w], NULL);
  const int err = atomic_load(&sched_error);
  const uint64_t lv = atomic_load(&tot_leaves), mc = atomic_load(&tot_machines), st = atomic_load(&tot_steps);
  if (err) {
    printf("error %d leaves %llu/%llu gives %llu takes %llu\n", err, (unsigned long long)lv,
           (unsigned long long)want_lv, (unsigned long long)atomic_load(&tot_gives),
           (unsigned long long)atomic_load(&tot_takes));
    return 1;
  }
  const int same = lv == want_lv && mc == want_mc && st == want_st;
  printf("%s %llu %llu %llu %llu %llu", same ? "ok" : "mismatch", (unsigned long long)lv, (unsigned long long)mc,
         (unsigned long long)st, (unsigned long long)atomic_load(&tot_gives),
         (unsigned long long)atomic_load(&tot_takes));
  if (!same)
    printf(" want %llu %llu %llu", (unsigned long long)want_lv, (unsigned long long)want_mc,
           (unsigned long long)want_st);
  printf("\n");
  return same ? 0 : 1;
}
