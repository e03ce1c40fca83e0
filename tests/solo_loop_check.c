/* solo_loop_check.c -- CPU check of the lone-lane step loop's accelerations.
 *
 * The register loop of the HIP kernels (run_solo in
 * erdos-machine-13_amd/csrc/e13_kernels.hip) advances a lone machine by
 * sweeps (runs of equal cells under a rule that rewrites the symbol it reads
 * and keeps the state) and by memoised block macro-steps (the run inside an
 * aligned 16-cell block, to the step that leaves it or for at most STAY
 * steps).  Both must give exactly what plain stepping gives: the same steps,
 * the same exit (a stage-gate candidate or an exceptional rule before the
 * step, the head leaving the window, which is not counted, or the head going
 * past a short lane tape of tw words, which is counted), the same tape, head,
 * state and ones count.  This program restates that loop in C (the
 * kernel's control flow and encodings, scalar) and compares it with a plain
 * one-step-at-a-time loop on random machines, tapes, heads, gate registers and
 * round lengths, biased towards the hard cases: few states (loops inside a
 * block), heads near the counter/block boundary at q 61 and near the window
 * edges, ones counts at or near the target.
 *
 * Usage: solo_loop_check SEED RUNS; prints "runs=N mismatches=M stays=S" and
 * exits 1 on a mismatch.  Built with -DPLANT=1 (ones mask of the boundary
 * block), -DPLANT=2 (gate margin 8 instead of 16), -DPLANT=4 (a word of a run
 * that crosses words counted as 33 cells), -DPLANT=3 (stay position
 * off by one) or -DPLANT=5 (a sweep allowed one cell past a short tape) it
 * carries a deliberate slip, which the check must catch.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef PLANT
#define PLANT 0
#endif
#define QMAX 1361 /* last live head position (q = cell + 63) */
#define QBLK 61   /* first q of the block region (cells >= -2), where ones are counted */
#define STAY 495  /* longest memoised in-block run */

static uint64_t splitmix(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

typedef struct {
  uint32_t tape[44];
  int q, st, ones;
} Machine;

/* rule rows as in LDS: two 16-bit entries write | dir << 1 | next << 8, or
   0x0F00 | kind << 12 for a halt / unbound cell */
static uint32_t rows[16];
static int target, qA, qR, st0r, nt1;
static int qlim = QMAX, tw = 43; /* the lane tape: cells q <= qlim = 32 * tw - 1 (QMAX for the whole window) */

static uint32_t entry(int st, uint32_t sym) { return (rows[st & 15] >> (sym << 4)) & 0xFFFFu; }

/* The plain loop: up to n steps; why = 1 exit before a step, 2 left the window,
   3 went past the lane tape (that step counts). */
static int plain(Machine *m, int n, int *why) {
  *why = 0;
  for (int r = 0; r < n; ++r) {
    const uint32_t qb = (uint32_t)m->q & 31u, sym = (m->tape[m->q >> 5] >> qb) & 1u;
    const uint32_t e = entry(m->st, sym);
    if ((e & 0x0F00u) == 0x0F00u ||
        (m->ones == target && (nt1 || (m->st == st0r && (m->q == qA || m->q == qR))))) {
      *why = 1;
      return r;
    }
    const uint32_t w = e & 1u;
    m->tape[m->q >> 5] ^= (w ^ sym) << qb;
    if (m->q >= QBLK) m->ones += (int)w - (int)sym;
    m->q += (int)(e & 2u) - 1;
    if ((unsigned)m->q > QMAX) {
      *why = 2;
      return r;
    }
    m->st = (int)((e >> 8) & 15u);
    if (m->q > qlim) {
      *why = 3;
      return r + 1;
    }
  }
  return n;
}

/* The memo: 1024 slots of (tag, value) in 16 x 64 words each, column `me`
   (the running lane's own) remapped to me ^ 32, cleared per call. */
static uint32_t mtag[16][64], mval[16][64];
static long stays;

/* The kernel's loop (scalar restatement): up to n steps. */
static int fast(Machine *m, int n, int *why, int me) {
  for (int w = 0; w < 16; w++)
    for (int l = 0; l < 64; l++)
      if (l != me) mtag[w][l] = 0;
  uint32_t sw0 = 0, sw1 = 0; /* sweep rules, bit = state, per symbol read */
  for (int k = 0; k < 16; ++k) {
    const uint32_t e0 = rows[k] & 0xFFFFu, e1 = rows[k] >> 16;
    if ((e0 & 0x0F00u) == ((uint32_t)k << 8) && (e0 & 1u) == 0u) sw0 |= 1u << k;
    if ((e1 & 0x0F00u) == ((uint32_t)k << 8) && (e1 & 1u) == 1u) sw1 |= 1u << k;
  }
  int q = m->q, st = m->st, ones = m->ones, wi = q >> 5;
  uint32_t word = m->tape[wi];
  int r = 0;
  *why = 0;
  for (; r < n; ++r) {
    const uint32_t qb = (uint32_t)q & 31u, sym = (word >> qb) & 1u;
    const uint32_t e = entry(st, sym), nx = e & 0x0F00u;
    if ((((sym ? sw1 : sw0) >> st) & 1u) != 0u) { /* a sweep */
      const uint32_t diff = word ^ (0u - sym);
      int k, to_end;
      if (e & 2u) {
        const uint32_t up = diff >> qb;
        to_end = up == 0u;
        k = up ? __builtin_ctz(up) : 32 - (int)qb;
      } else {
        const uint32_t dn = diff << (31u - qb);
        to_end = dn == 0u;
        k = dn ? __builtin_clz(dn) : (int)qb + 1;
      }
      if (to_end) { /* the run goes on through the following words while they are all sym */
        const uint32_t fill = 0u - sym;
        int wj = wi + ((e & 2u) ? 1 : -1);
        while (wj >= 0 && wj < tw && m->tape[wj] == fill) {
          k += PLANT == 4 ? 33 : 32;
          wj += (e & 2u) ? 1 : -1;
        }
        const uint32_t wd = (wj >= 0 && wj < tw ? m->tape[wj] : ~fill) ^ fill;
        k += (e & 2u) ? __builtin_ctz(wd) : __builtin_clz(wd);
      }
      if (e & 2u) {
        const int lim = qlim - q + (PLANT == 5);
        if (k > lim) k = lim;
      } else {
        if (k > q) k = q;
      }
      if (k > n - r) k = n - r;
      if (ones == target) {
        if (nt1 | ((st == st0r) & ((q == qA) | (q == qR)))) k = 0;
        else if (st == st0r) {
          if (e & 2u) {
            if (qA > q && qA - q < k) k = qA - q;
            if (qR > q && qR - q < k) k = qR - q;
          } else {
            if (qA < q && q - qA < k) k = q - qA;
            if (qR < q && q - qR < k) k = q - qR;
          }
        }
      }
      if (k > 0) {
        q += (e & 2u) ? k : -k;
        r += k - 1;
        if ((q >> 5) != wi) {
          m->tape[wi] = word;
          wi = q >> 5;
          word = m->tape[wi];
        }
        continue;
      }
    }
    const int qn = q & ~15;
    const int margin = PLANT == 2 ? 8 : 16;
    const int far = (ones - target > margin) | (target - ones > margin);
    const int no_gate_cell = !nt1 & ((qA >> 4) != (qn >> 4)) & ((qR >> 4) != (qn >> 4));
    if ((qn >= 16) & (qn + 16 <= qlim) & (far | no_gate_cell)) { /* a block macro-step */
      const uint32_t nb = qb & ~15u, byt = (word >> nb) & 0xFFFFu, pos = qb & 15u;
      const uint32_t key = ((uint32_t)st << 20) | (byt << 4) | pos;
      const uint32_t slot = (key * 0x9E3779B1u) >> 22;
      int col = (int)(slot & 63u);
      if (col == me) col ^= 32;
      const int ws = (int)(slot >> 6);
      const uint32_t tag = 0x80000000u | key;
      uint32_t val;
      if (mtag[ws][col] != tag) {
        uint32_t v = byt, s2 = (uint32_t)st;
        int p = (int)pos, k2 = 0, right = 0, code = 0;
        while (k2 < STAY) {
          const uint32_t b = (v >> p) & 1u, e2 = entry((int)s2, b);
          if ((e2 & 0x0F00u) == 0x0F00u) break;
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
        if (k2 == STAY && code == 0) code = STAY + 1 + p;
        val = v | (s2 << 16) | ((uint32_t)right << 20) | ((uint32_t)code << 21);
        mtag[ws][col] = tag;
        mval[ws][col] = val;
      } else {
        val = mval[ws][col];
      }
      const int code = (int)(val >> 21), k3 = code < STAY ? code : STAY;
      if ((code > 0) & (k3 <= n - r)) {
        const uint32_t nv = val & 0xFFFFu;
        word = (word & ~(0xFFFFu << nb)) | (nv << nb);
        const uint32_t cm = PLANT == 1 ? (qn + 16 > QBLK ? 0xFFFFu : 0u)
                                       : (qn >= QBLK ? 0xFFFFu
                                                     : (qn + 16 > QBLK ? (0xFFFFu << (QBLK - qn)) & 0xFFFFu : 0u));
        ones += __builtin_popcount(nv & cm) - __builtin_popcount(byt & cm);
        if (code > STAY) ++stays;
        q = code > STAY ? qn + (code - STAY - 1) + (PLANT == 3) : ((val >> 20) & 1u) ? qn + 16 : qn - 1;
        st = (int)((val >> 16) & 15u);
        r += k3 - 1;
        if ((q >> 5) != wi) {
          m->tape[wi] = word;
          wi = q >> 5;
          word = m->tape[wi];
        }
        continue;
      }
    }
    if ((nx == 0x0F00u) | ((ones == target) & (nt1 | ((st == st0r) & ((q == qA) | (q == qR)))))) {
      *why = 1; /* the kernel hands this step to step_once */
      break;
    }
    const uint32_t w = e & 1u;
    word ^= (w ^ sym) << qb;
    ones += q >= QBLK ? (int)w - (int)sym : 0;
    q += (int)(e & 2u) - 1;
    if ((unsigned)q > (unsigned)QMAX) {
      *why = 2;
      break;
    }
    st = (int)(nx >> 8);
    if (q > qlim) { /* past the lane tape: the step counts, the kernel pauses the machine */
      r += 1;
      *why = 3;
      break;
    }
    if ((q >> 5) != wi) {
      m->tape[wi] = word;
      wi = q >> 5;
      word = m->tape[wi];
    }
  }
  m->tape[wi] = word; /* the word of the last cell inside the window */
  m->q = q;
  m->st = st;
  m->ones = ones;
  return r;
}

int main(int argc, char **argv) {
  uint64_t s = argc > 1 ? strtoull(argv[1], 0, 0) : 12345;
  const long runs = argc > 2 ? atol(argv[2]) : 400000;
  long bad = 0;
  for (long it = 0; it < runs; it++) {
    const int nq = (it % 3 == 0) ? 2 + (int)(splitmix(&s) % 3) : 2 + (int)(splitmix(&s) % 13);
    for (int k = 0; k < 16; k++) {
      uint32_t pr = 0;
      for (int b = 0; b < 2; b++) {
        const uint32_t v = (uint32_t)(splitmix(&s) % 100);
        uint32_t e;
        if (k >= nq || v < 3) e = 0x0F00u | (1u + (uint32_t)(splitmix(&s) % 3)) << 12;
        else
          e = (uint32_t)(splitmix(&s) & 1u) | (uint32_t)((splitmix(&s) & 1u) << 1) |
              ((uint32_t)(splitmix(&s) % (uint64_t)nq) << 8);
        pr |= e << (16 * b);
      }
      rows[k] = pr;
    }
    Machine a;
    memset(&a, 0, sizeof a);
    for (int w = 0; w < 43; w++) a.tape[w] = (uint32_t)splitmix(&s) & (uint32_t)splitmix(&s);
    if (it % 4 == 3) /* long runs: whole words of zeros or of ones, so sweeps cross words */
      for (int w = 0; w < 43; w++) {
        const uint64_t x = splitmix(&s) % 4;
        if (x) a.tape[w] = x == 1 ? ~0u : 0u;
      }
    a.tape[42] &= (1u << (QMAX + 1 - 1344)) - 1u;
    /* half the runs on a short lane tape of tw words: the cells past it blank */
    tw = it % 2 ? 18 + (int)(splitmix(&s) % 25) : 43;
    qlim = tw < 43 ? 32 * tw - 1 : QMAX;
    for (int w = tw; w < 43; w++) a.tape[w] = 0u;
    const int where = (int)(it % 8);
    a.q = where == 1 ? 52 + (int)(splitmix(&s) % 20)
          : where == 2 ? (int)(splitmix(&s) % 24)
          : where == 3 ? qlim - (int)(splitmix(&s) % 24)
                       : (int)(splitmix(&s) % (qlim + 1));
    a.st = (int)(splitmix(&s) % (uint64_t)nq);
    for (int q = QBLK; q <= QMAX; q++) a.ones += (int)((a.tape[q >> 5] >> (q & 31)) & 1u);
    const int nt = 1 + (int)(splitmix(&s) % 9), tm = (int)(splitmix(&s) % 4);
    target = tm == 0 ? a.ones + (int)(splitmix(&s) % 19) - 9 : tm == 1 ? a.ones : (1 << nt);
    nt1 = nt == 1 && splitmix(&s) % 2;
    st0r = (int)(splitmix(&s) % (uint64_t)nq);
    qA = splitmix(&s) % 3 == 0 ? -1000 : (int)((a.q + (int)(splitmix(&s) % 200) - 100 + QMAX + 1) % (QMAX + 1));
    qR = splitmix(&s) % 3 == 0 ? -1000 : (int)((a.q + (int)(splitmix(&s) % 200) - 100 + QMAX + 1) % (QMAX + 1));
    const int n = 1 + (int)(splitmix(&s) % 5000), me = (int)(splitmix(&s) % 64);
    Machine b = a;
    int wa, wb;
    const int ra = plain(&a, n, &wa), rb = fast(&b, n, &wb, me);
    if (ra != rb || wa != wb || a.q != b.q || (wa != 2 && a.st != b.st) || a.ones != b.ones ||
        memcmp(a.tape, b.tape, sizeof a.tape)) {
      if (bad < 3)
        printf("mismatch run=%ld n=%d: plain r=%d why=%d q=%d st=%d ones=%d | loop r=%d why=%d q=%d st=%d ones=%d\n",
               it, n, ra, wa, a.q, a.st, a.ones, rb, wb, b.q, b.st, b.ones);
      bad++;
    }
  }
  printf("runs=%ld mismatches=%ld stays=%ld\n", runs, bad, stays);
  return bad ? 1 : 0;
}
