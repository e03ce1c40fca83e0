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
    fprintf(stderr, "usage: e13ref_dump keys|hash ...\n");
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
  (void)e13ref_main;
  fprintf(stderr, "unknown mode %s\n", argv[1]);
  return 1;
}
