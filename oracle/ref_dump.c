/* ref_dump.c -- test infrastructure: a harness around the reference program.
 *
 * Built by oracle/Makefile with -DREF_SRC naming the reference's own
 * src/e13search.c, which it includes where it lies (nothing is copied).  The
 * reference's main() is renamed; this main() fills the reference's KEY/TGT
 * exactly as its main does (e13search.c:437-444) and then uses the
 * reference's own static run_full()/run() to print what the reference keeps
 * but never prints:
 *
 *   e13ref_dump keys            KEY[0..1363] and TGT[1..9], one hex per line
 *   e13ref_dump hash < lines    for each SURVIVOR line, the counter hashes
 *                               ch[0..2] left by run() on that table with
 *                               every cell fixed ('?---' read as halt)
 *   e13ref_dump probe machine X FMAX LO HI SEED N [CELL...]
 *                               the per-probe sums (leaves, nodes, steps,
 *                               survivors; hex of the doubles) of the
 *                               reference's own probe() for probes 0..N-1 of
 *                               the estimate problem (argv as in its estimate
 *                               mode), probe i drawing from the splitmix state
 *                               splitmix(0x5EED0000 + (SEED << 40) + i)
 *   e13ref_dump setup machine X FMAX LO HI [CELL...]
 *                               what the reference's set_sym() + setup()
 *                               leave for that problem (argv as in its
 *                               estimate mode: LO < 0 no set_sym, cells give
 *                               free mode): one line with the budget, SYM_LO,
 *                               SYM_HI, SYM_N0 and the table codes; setup()'s
 *                               own message and exit status 1 when it refuses
 */
#define main e13ref_main
#include REF_SRC
#undef main

static void ref_init(void) {
  uint64_t seed = 20261028;
  for (int i = 0; i < TAPE; i++) KEY[i] = splitmix(&seed);
  for (int n = 1; n <= NS; n++) {
    TGT[n] = 0;
    for (int c = -1; c <= (1 << n) - 2; c++) TGT[n] ^= KEY[c + LO];
  }
}

static int parse_tab(const char *tb, int *tab) {
  int n = 0;
  const char *p = tb;
  while (*p && *p != '\n' && n < 2 * MAXQ) {
    while (*p == ' ' || *p == '|') p++;
    if (!*p || *p == '\n') break;
    if (*p == '?') p++;
    char cell[16];
    int k = 0;
    while (*p && *p != ' ' && *p != '|' && *p != '\n' && k < 15) cell[k++] = *p++;
    cell[k] = 0;
    tab[n++] = parse_cell(cell);
  }
  return n;
}

int main(int argc, char **argv) {
  ref_init();
  if (argc < 2) {
    fprintf(stderr, "usage: e13ref_dump keys|hash|probe|setup ...\n");
    return 1;
  }
  if (!strcmp(argv[1], "keys")) {
    for (int i = 0; i < TAPE; i++) printf("%016llx\n", (unsigned long long)KEY[i]);
    for (int n = 1; n <= NS; n++) printf("%016llx\n", (unsigned long long)TGT[n]);
    return 0;
  }
  if (!strcmp(argv[1], "hash")) {
    char line[8192];
    Sim *s = (Sim *)malloc(sizeof(Sim));
    while (fgets(line, sizeof line, stdin)) {
      char *tb = strstr(line, "table=");
      if (!tb) continue;
      int tab[2 * MAXQ], X = 0;
      for (int i = 0; i < 2 * MAXQ; i++) tab[i] = -1;
      char *xp = strstr(line, "X=");
      if (xp) X = atoi(xp + 2);
      parse_tab(tb + 6, tab);
      int r = run_full(tab, s);
      printf("HASH X=%d ch0=%016llx ch1=%016llx ch2=%016llx result=%d steps=%ld\n", X, (unsigned long long)s->ch[0],
             (unsigned long long)s->ch[1], (unsigned long long)s->ch[2], r, s->t);
    }
    free(s);
    return 0;
  }
  if (!strcmp(argv[1], "probe") && argc >= 9) {
    load(argv[2]);
    int X = atoi(argv[3]), fmax = atoi(argv[4]), tab[2 * MAXQ];
    unsigned long long seed = strtoull(argv[7], NULL, 10);
    long n = atol(argv[8]);
    unsigned char freebase[2 * MAXQ] = {0};
    if (atoi(argv[5]) >= 0) set_sym(X, atoi(argv[5]), atoi(argv[6]));
    for (int i = 9; i < argc; i++) freebase[2 * (argv[i][0] - 'A') + (argv[i][1] - '0')] = 1;
    int budget = setup(X, fmax, argc > 9 ? freebase : NULL, tab);
    if (budget < 0) return 1;
    Ctx *x = (Ctx *)calloc(1, sizeof(Ctx));
    x->nq = NQ - (X >= 0);
    printf("probe X=%d budget=%d seed=%llu n=%ld\n", X, budget, seed, n);
    for (long i = 0; i < n; i++) {
      uint64_t st = 0x5EED0000ULL + ((uint64_t)seed << 40) + (uint64_t)i;
      uint64_t rng = splitmix(&st);
      double o[4] = {0, 0, 0, 0};
      probe(x, tab, budget, &rng, o);
      uint64_t b[4];
      memcpy(b, o, sizeof b);
      printf("%ld %016llx %016llx %016llx %016llx\n", i, (unsigned long long)b[0], (unsigned long long)b[1],
             (unsigned long long)b[2], (unsigned long long)b[3]);
    }
    free(x);
    return 0;
  }
  if (!strcmp(argv[1], "setup") && argc >= 7) {
    load(argv[2]);
    int X = atoi(argv[3]), fmax = atoi(argv[4]), tab[2 * MAXQ];
    unsigned char freebase[2 * MAXQ] = {0};
    if (atoi(argv[5]) >= 0) set_sym(X, atoi(argv[5]), atoi(argv[6]));
    for (int i = 7; i < argc; i++) freebase[2 * (argv[i][0] - 'A') + (argv[i][1] - '0')] = 1;
    int budget = setup(X, fmax, argc > 7 ? freebase : NULL, tab); /* exits 1 on a symmetry refusal */
    int nq = NQ - (X >= 0);
    printf("setup budget=%d sym_lo=%d sym_hi=%d sym_n0=%d tab=", budget, SYM_LO, SYM_HI, SYM_N0);
    for (int i = 0; i < 2 * nq; i++) printf("%s%d", i ? "," : "", tab[i]);
    printf("\n");
    return 0;
  }
  (void)e13ref_main;
  fprintf(stderr, "unknown mode %s\n", argv[1]);
  return 1;
}
