// smdp-gen: writes the seeded synthetic graphs as DIMACS9 .gr files so the
// gpu-* drivers can read them with -f 0.
//   smdp-gen uniform <n> <deg> <wlo> <whi> <seed> <out.gr>
//   smdp-gen rmat <scale> <ef> <wlo> <whi> <seed> <out.gr> [--sym] [--no-perm]
//   smdp-gen tree <n> <seed> <out.gr>        (Zipf(1.5) child counts, cap 4096)
#include <stdlib.h>
#include <string.h>

#include "cli_common.h"

static int usage() {
  fprintf(stderr,
          "usage: smdp-gen uniform <n> <deg> <wlo> <whi> <seed> <out.gr>\n"
          "       smdp-gen rmat <scale> <ef> <wlo> <whi> <seed> <out.gr> [--sym] [--no-perm]\n"
          "       smdp-gen tree <n> <seed> <out.gr>\n");
  return 2;
}

int main(int argc, char** argv) {
  if (argc < 2) return usage();
  smdp_hgraph* g = nullptr;
  const char* out = nullptr;
  int rc;
  if (!strcmp(argv[1], "uniform") && argc == 8) {
    rc = smdp_gen_uniform(atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), atoi(argv[5]), strtoull(argv[6], 0, 10), &g);
    out = argv[7];
  } else if (!strcmp(argv[1], "rmat") && argc >= 8) {
    int sym = 0, perm = 1;
    for (int i = 8; i < argc; ++i) {
      if (!strcmp(argv[i], "--sym")) sym = 1;
      if (!strcmp(argv[i], "--no-perm")) perm = 0;
    }
    rc = smdp_gen_rmat(atoi(argv[2]), atoi(argv[3]), 0.57, 0.19, 0.19, atoi(argv[4]), atoi(argv[5]), sym, perm,
                       strtoull(argv[6], 0, 10), &g);
    out = argv[7];
  } else if (!strcmp(argv[1], "tree") && argc == 5) {
    rc = smdp_gen_zipf_tree(atoi(argv[2]), 1.5, 4096, strtoull(argv[3], 0, 10), &g);
    out = argv[4];
  } else {
    return usage();
  }
  if (rc != SMDP_OK) return cli_fail("generating", rc);
  rc = smdp_write_dimacs9(out, g);
  if (rc != SMDP_OK) return cli_fail("writing", rc);
  fprintf(stderr, "wrote %s: %d vertices, %d arcs, source %d\n", out, g->n, g->m, g->source + 1);
  smdp_hgraph_free(g);
  return 0;
}
