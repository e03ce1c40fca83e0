// e13_deepk.hip -- the deep stage check (src/deep.c:15-84) and the counter
// re-check (src/recheck.c:22-104) of survivors on gfx950.
//
// One wave per survivor: the survivor's whole tape (left + 2^nmax + 4096
// cells, 5 KB at nmax = 15) sits in LDS as plain words (cell k at bit k + left
// of word (k + left) >> 5), and the wave runs the machine in wave-uniform
// control flow, so every value of the loop lives in scalar registers.  Two
// things make a long run cheap:
//  - the current tape word is cached in a register and written back only
//    when the head crosses into another word;
//  - sweeps: under a rule that rewrites the symbol it reads and keeps the
//    state, a run of equal cells is one step per cell that moves only the
//    head, so the whole run is one iteration (ctz/clz of the cached word, and
//    past its end a read of the next 64 words, one per lane, with a ballot).
//    A stage-start candidate cell ahead, the window's edge and the step limit
//    cut a sweep short, so every check of the reference's loop still sees
//    every configuration it would see.
// The reference's order inside a step (deep.c:51-71, recheck.c:64-79): stage
// start, step limit, window, read (free cell / halt), write, move, state, t++.
// Those tests run in full only where one can fire; between, a tight loop in
// 32-bit counts makes the steps and sweeps (a lone wave issues about one
// scalar instruction per 5 clocks, so the long survivors that end a deep check
// alone are bound by the length of this loop).
// A launch runs at most `budget` iterations per survivor, then saves the tape
// and the registers; the host launches again until every survivor is done.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "e13_dev.h"

namespace {

typedef __attribute__((address_space(3))) uint32_t lds_word_t;
__device__ __forceinline__ uint32_t ldw(int w) { return *(const lds_word_t *)(size_t)(uint32_t)(w * 4); }
__device__ __forceinline__ void stw(int w, uint32_t v) { *(lds_word_t *)(size_t)(uint32_t)(w * 4) = v; }
__device__ __forceinline__ int U(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint32_t UW(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ long long U64(long long v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((unsigned long long)v >> 32));
  return (long long)(((unsigned long long)hi << 32) | lo);
}

// A rule entry (16 bits): write | right << 1 | next << 8 (5 bits), or
// 0x8000 | kind << 13 for a cell that is not a move (kind 1 halt, 2 a free cell
// the search never read).
#define DE_STOP 0x8000u

// Cells -1 .. 2^n - 2 (bits left - 1 .. left + 2^n - 2) all ones: the block of
// a stage start (deep.c:52-55).  The cached word is in LDS.
__device__ __forceinline__ bool deep_block(int lane, int left, int n) {
  const int qlo = left - 1, qhi = left + (1 << n) - 2;
  bool ok = true;
  for (int w0 = qlo >> 5; w0 <= (qhi >> 5); w0 += 64) {
    const int w = w0 + lane;
    bool good = true;
    if (w <= (qhi >> 5)) {
      uint32_t m = 0xFFFFFFFFu;
      if (w == (qlo >> 5)) m &= 0xFFFFFFFFu << (qlo & 31);
      if (w == (qhi >> 5)) m &= 0xFFFFFFFFu >> (31 - (qhi & 31));
      good = (ldw(w) & m) == m;
    }
    if (__ballot(!good) != 0ull) {
      ok = false;
      break;
    }
  }
  return ok;
}

} // namespace

template <int MODE>
__global__ void __launch_bounds__(64) k_deep(E13DeepArgs a) {
  extern __shared__ __align__(16) uint32_t smem[];
  const int lane = (int)threadIdx.x;
  const int id = U(a.list[blockIdx.x]);
  E13DeepState *S = a.state + id;
  if ((uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void *)smem != 0u) {
    // the LDS addressing needs the dynamic area at address 0
    if (lane == 0) S->status = 3;
    return;
  }
  const int W = a.W, left = a.left, nmax = a.nmax;
  uint32_t *gt = a.tape + (size_t)id * (size_t)W;
  for (int w = lane; w < W; w += 64) stw(w, gt[w]);
  const uint32_t rowv = lane < E13_DEEP_MAXQ ? a.rows[(size_t)id * E13_DEEP_MAXQ + lane] : 0u;
  const int st0 = U(a.param[4 * id + 0]), h0q = U(a.param[4 * id + 1]), rel = U(a.param[4 * id + 2]);
  // sweep rules: (state, symbol) whose move keeps the state and rewrites the symbol read
  const uint32_t e0v = rowv & 0xFFFFu, e1v = rowv >> 16;
  const bool mine = lane < E13_DEEP_MAXQ;
  const uint32_t sw0 = (uint32_t)__ballot(mine && !(e0v & DE_STOP) && ((e0v >> 8) & 31u) == (uint32_t)lane && !(e0v & 1u));
  const uint32_t sw1 = (uint32_t)__ballot(mine && !(e1v & DE_STOP) && ((e1v >> 8) & 31u) == (uint32_t)lane && (e1v & 1u));
  int q = U(S->q), st = U(S->st), ones = U(S->ones), n = U(S->n), minq = U(S->minq);
  long long t = U64(S->t), limit = U64(S->limit);
  int status = 0, fail = 0, fail_q = 0, fail_st = 0, fail_sym = 0;
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  // the stage sought: its block, the head of its start, the window's right edge
  int blk = 1 << n;
  int at = rel ? h0q + blk - 2 : h0q;
  const int lo = MODE == 0 ? 2 : 1; // deep.c:57 head < -left + 2; the re-check: the tape held here
  int hi = MODE == 0 ? left + blk + 2000 : 32 * W - 2;
  const int qb0 = left - 2; // ones are counted at cells >= -2
  int wi = q >> 5;
  uint32_t word = UW(ldw(wi));
  for (long long it = 0; it < a.budget;) {
    // ---- the reference's tests on this configuration, in its order (deep.c:51-60,
    // recheck.c:64-72) ----
    // stage start (deep.c:51-56)
    if ((ones == blk) & (st == st0) & (q == at)) {
      stw(wi, word);
      if (deep_block(lane, left, n)) {
        // the counter side of the tape at this start, for the host's counts
        uint32_t *sn = a.snap + ((size_t)id * (size_t)(nmax + 1) + (size_t)n) * (size_t)(left >> 5);
        for (int w = lane; w < (left >> 5); w += 64) sn[w] = ldw(w);
        ++n;
        if (n > nmax) {
          status = 1;
          break;
        }
        blk = 1 << n;
        limit = t + a.kmul * (1ll << (2 * n)) + 100000;
        at = rel ? h0q + blk - 2 : h0q;
        if (MODE == 0) hi = left + blk + 2000;
        ++it;
        continue;
      }
    }
    if (t >= limit) { // deep.c:56, recheck.c:69
      status = 1;
      fail = 1;
      break;
    }
    if ((q < lo) | (q > hi)) { // deep.c:57: off the tape; the re-check: past the tape held here
      status = MODE == 0 ? 1 : 2;
      fail = 2;
      fail_q = q;
      break;
    }
    {
      const uint32_t sym = (word >> ((uint32_t)q & 31u)) & 1u;
      const uint32_t e = ((uint32_t)__builtin_amdgcn_readlane((int)rowv, st) >> (sym << 4)) & 0xFFFFu;
      if (e & DE_STOP) { // deep.c:59-60: a free cell never read, or a halt
        status = 1;
        fail = ((e >> 13) & 3u) == 2u ? 3 : 4;
        fail_st = st;
        fail_sym = (int)sym;
        break;
      }
    }
    // ---- none fires here: steps and sweeps in a tight loop, in 32-bit counts,
    // back to the tests above on the stage-start cell, on a cell that stops the
    // run, and once `room` steps are made: the steps left before the step limit
    // and before the head could reach a window edge (it moves one cell a step),
    // so the limit and window tests pass on every configuration in between ----
    const int room = (int)min(min(limit - t, (long long)min(q - lo, hi - q) + 1), (long long)(1 << 30));
    const int iters = (int)min(a.budget - it, (long long)(1 << 30));
    int done = 0, nit = 0;
    for (;;) {
      const uint32_t qb = (uint32_t)q & 31u;
      const uint32_t sym = (word >> qb) & 1u;
      const uint32_t e = (uint32_t)__builtin_amdgcn_readlane((int)rowv, st) >> (sym << 4);
      if (e & DE_STOP) break; // (never on the first pass: tested above)
      int k = 0;
      uint32_t wv = 0u; // for a run past word wi: word wi + 1 + lane (right) or wi - 1 - lane (left)
      if ((((sym ? sw1 : sw0) >> st) & 1u) != 0u) {
        // a sweep: k cells of the run of `sym` cells from q in the direction of travel
        const uint32_t diff = word ^ (0u - sym);
        bool to_end;
        if (e & 2u) {
          const uint32_t up = diff >> qb;
          to_end = up == 0u;
          k = up ? __builtin_ctz(up) : 32 - (int)qb;
        } else {
          const uint32_t dn = diff << (31u - qb);
          to_end = dn == 0u;
          k = dn ? __builtin_clz(dn) : (int)qb + 1;
        }
        if (to_end) { // the run may go on: the next 64 words, one per lane
          stw(wi, word);
          const int wj = wi + ((e & 2u) ? 1 + lane : -1 - lane);
          const uint32_t fill = 0u - sym;
          wv = ~fill;
          if ((wj >= 0) & (wj < W)) wv = ldw(wj);
          const uint64_t mism = __ballot(wv != fill);
          if (mism) {
            const int j = __builtin_ctzll(mism);
            const uint32_t wd = (uint32_t)__builtin_amdgcn_readlane((int)wv, j) ^ fill;
            k += 32 * j + ((e & 2u) ? __builtin_ctz(wd) : __builtin_clz(wd));
          } else {
            k += 32 * 64;
          }
        }
        k = min(k, room - done); // inside the window and the step limit
        if ((ones == blk) & (st == st0)) { // stop on the start cell ahead: the tests above take it
          if ((e & 2u) && at > q) k = min(k, at - q);
          if (!(e & 2u) && at < q) k = min(k, q - at);
        }
      }
      if (k > 0) {
        q += (e & 2u) ? k : -k;
        done += k;
        minq = min(minq, q);
        if ((q >> 5) != wi) { // only a run that reached the end of word wi (written back above)
          const int wn = q >> 5;
          const int jn = (e & 2u) ? wn - wi - 1 : wi - wn - 1; // lane that read word wn
          wi = wn;
          // the sweep wrote nothing: the word is the one read for the run, unless
          // the run went on past all 64 of them
          word = jn < 64 ? (uint32_t)__builtin_amdgcn_readlane((int)wv, jn) : UW(ldw(wi));
        }
      } else {
        // one step (deep.c:62-66)
        const uint32_t w1 = e & 1u;
        word ^= (w1 ^ sym) << qb;
        if (q >= qb0) ones += (int)w1 - (int)sym;
        q += (e & 2u) ? 1 : -1;
        st = (int)((e >> 8) & 31u);
        ++done;
        minq = min(minq, q);
        if ((q >> 5) != wi) {
          stw(wi, word);
          wi = q >> 5;
          word = UW(ldw(wi));
        }
      }
      ++nit;
      if ((done >= room) | (nit >= iters) | (q == at)) break;
    }
    t += done;
    it += nit;
  }
  stw(wi, word);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  for (int w = lane; w < W; w += 64) gt[w] = ldw(w);
  if (lane == 0) {
    S->q = q;
    S->st = st;
    S->ones = ones;
    S->n = n;
    S->minq = minq;
    S->t = t;
    S->limit = limit;
    S->status = status;
    if (status) {
      S->fail = fail;
      S->fail_n = n;
      S->fail_t = t;
      S->fail_q = fail_q;
      S->fail_st = fail_st;
      S->fail_sym = fail_sym;
    }
  }
}

extern "C" {

int e13k_launch_deep(const E13DeepArgs *a, int nblocks, hipStream_t s) {
  const size_t lds = (size_t)a->W * 4;
  if (a->mode == 0) hipLaunchKernelGGL(k_deep<0>, dim3(nblocks), dim3(64), lds, s, *a);
  else hipLaunchKernelGGL(k_deep<1>, dim3(nblocks), dim3(64), lds, s, *a);
  return (int)hipGetLastError();
}

} // extern "C"
